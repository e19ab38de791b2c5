"""The nonlinear Galerkin path: Newton's method with the tangent re-assembled
on the GPU every step.

Mirrors the nonlinear branch of ``solve_galerkin``
(``src/jaxiga/methods/galerkin.py:337-355``) and ``newton``
(``src/jaxiga/solvers/newton.py:27-145``): the residual ``r(u)`` is the
gradient of the total energy on the free dofs, each step solves ``K(u) du =
-r`` with the linear options, and a backtracking line search halves the step
(at most 20 times) until the residual norm decreases; the iteration stops at
``||r|| <= tol`` or after ``max_steps``, and ``stats`` carries the
reference's keys (``iterations, residual, converged, finite,
iteration_limit, stalled, history``).

The reference obtains residual and tangent by autodiff of the energy; the
element kernel here (``iga_neohooke``) writes them out for the neo-Hookean
energy. What is fixed across steps is built once: the basis data at the
Gauss points, the body force at the points, the Neumann load, the CSR
pattern and the element-to-CSR gather; each step re-runs only the element
kernel and the gather pass of ``iga_assemble`` (``passes = 2``), which sums
the new element tangents into the same pattern.
"""

from __future__ import annotations

import math
import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd.evaluation import evaluate
from jaxiga_amd.pointset import gauss
from jaxiga_amd.runtime import lib, stream

MAX_BACKTRACKS = 20


class NeoHookeanOperator:
    """Residual and tangent of a neo-Hookean problem on the free dofs."""

    def __init__(self, ctx, params):
        import torch

        from jaxiga_amd.methods import call_points, element_workspace, neumann_load

        problem = ctx.problem
        space = ctx.space
        if problem.kernel != "neohooke":
            raise TypeError("NeoHookeanOperator needs a Hyperelasticity problem")
        self.ctx = ctx
        self.mu, self.lam = problem.lame_parameters(params)
        self.basis = evaluate(space, gauss(space, ctx.n_q1), outputs=("R", "dR", "x", "w"))
        dev = ctx.f2f_d.device
        self.edofs = torch.as_tensor(np.ascontiguousarray(space.elem_dofs, dtype=np.int32), device=dev)
        self.body = None
        if problem.source is not None:
            src = problem.source
            if callable(src):
                body = call_points(src, self.basis.x, params=params, vec=space.vec)
            else:
                body = torch.as_tensor(np.broadcast_to(np.asarray(src, float), (space.vec,)),
                                       dtype=torch.float64, device=dev).expand(*self.basis.w.shape, space.vec)
            self.body = body.detach().to(torch.float64).contiguous()
        self.ext = torch.zeros(ctx.n_free, dtype=torch.float64, device=dev)
        if problem.neumann:
            neumann_load(ctx, params, self.ext)
        self.pattern = ctx.assembly
        self.ke = element_workspace(ctx, _lib.IGA_FP64)
        # the gather pass of iga_assemble, over the element tangents in ke
        args = _lib.IgaAsmArgs()
        args.problem, args.vec, args.precision = _lib.IGA_ELASTICITY, space.vec, _lib.IGA_FP64
        args.full_to_free = ctx.f2f_d.data_ptr()
        args.freed = ctx.free_d.data_ptr() if ctx.n_free else None
        args.n_free, args.batch = ctx.n_free, 1
        pat = self.pattern
        args.nnz, args.rows, args.cols = pat.nnz, pat.rows_d.data_ptr(), pat.cols_d.data_ptr()
        args.indptr = pat.indptr_d.data_ptr()
        args.ke_work, args.ke_work_bytes = self.ke.data_ptr(), self.ke.numel()
        args.passes = 2
        args.ke_full = 1  # k_neohooke writes the tangents as full rows
        self.args = args
        self.u_full = torch.zeros(space.n_dofs, dtype=torch.float64, device=dev)
        self.presc = torch.as_tensor(ctx.dofmap.prescribed, dtype=torch.long, device=dev)
        self.free = ctx.free_d.long()
        if len(ctx.dofmap.prescribed):
            self.u_full[self.presc] = torch.as_tensor(ctx.dofmap.values, dtype=torch.float64, device=dev)

    def _launch(self, u_free, rfree, ke):
        self.u_full[self.free] = u_free
        b = self.basis
        _lib.check(lib().iga_neohooke(b.w.shape[0], b.w.shape[1], b.R.shape[-1], self.ctx.space.dim,
                                      _lib.ptr(b.R), _lib.ptr(b.dR), _lib.ptr(b.w), _lib.ptr(self.edofs),
                                      _lib.ptr(self.u_full), _lib.ptr(self.ctx.f2f_d), self.mu, self.lam,
                                      _lib.ptr(self.body), _lib.ptr(rfree), _lib.ptr(ke), stream()),
                   "iga_neohooke")

    def residual(self, u_free):
        """``r = grad Pi`` on the free dofs (internal minus body and Neumann forces)."""
        r = -self.ext.clone()
        self._launch(u_free, r, None)
        return r

    def tangent(self, u_free):
        """FP64 CSR values of the tangent at ``u_free`` (the pattern of ``ctx.assembly``)."""
        import torch

        vals = torch.empty(self.pattern.nnz, dtype=torch.float64, device=u_free.device)
        self._launch(u_free, None, self.ke)
        self.args.vals = vals.data_ptr()
        _lib.check(lib().iga_assemble(self.ctx.space.device().struct, self.ctx.rule(), self.args, stream()),
                   "iga_assemble")
        return vals


def newton(op: NeoHookeanOperator, linear, u0=None, tol=1e-8, max_steps=25, line_search=True):
    """Newton with backtracking on the residual norm; ``(u_free, stats)``."""
    import torch

    n = op.ctx.n_free
    u = torch.zeros(n, dtype=torch.float64, device=op.ext.device) if u0 is None else u0.clone()

    def norm(v):
        return float(torch.linalg.vector_norm(op.residual(v)))

    r_norm = norm(u)
    hist = [r_norm]
    step, alpha = 0, 1.0
    while math.isfinite(r_norm) and r_norm > tol and step < max_steps and alpha > 0:
        r = op.residual(u)
        du, _ = LIN.sparse_solve(linear, op.pattern, op.tangent(u), -r)
        alpha = 1.0
        if line_search:
            trial, k = norm(u + du), 0
            while (not math.isfinite(trial) or trial >= r_norm) and k < MAX_BACKTRACKS:
                alpha *= 0.5
                trial = norm(u + alpha * du)
                k += 1
            alpha = alpha if (math.isfinite(trial) and trial < r_norm) else 0.0
        if alpha > 0:
            u = u + alpha * du
            r_norm = norm(u)
        step += 1
        hist.append(r_norm)
    converged = r_norm <= tol
    stalled = alpha <= 0 and not converged
    r_norm = norm(u)
    finite = math.isfinite(r_norm) and bool(torch.isfinite(u).all())
    return u, {"iterations": step, "residual": r_norm, "converged": finite and r_norm <= tol,
               "finite": finite, "iteration_limit": step >= max_steps and not converged,
               "stalled": stalled, "history": hist}


def solve_nonlinear(ctx, params, linear, initial_guess=None, tol=1e-8, max_steps=25, line_search=True):
    """The nonlinear branch of the Galerkin driver; ``(u_full, stats)``."""
    import torch

    t0 = time.perf_counter()
    op = NeoHookeanOperator(ctx, params)
    u0 = None
    if initial_guess is not None:
        u0 = torch.as_tensor(initial_guess, dtype=torch.float64, device=op.ext.device)
        if u0.shape[0] == ctx.space.n_dofs:
            u0 = u0[op.free]
    u, stats = newton(op, linear, u0, tol, max_steps, line_search)
    if not stats["converged"]:
        import warnings

        warnings.warn(f"Newton did not converge (residual {stats['residual']:.3e} after "
                      f"{stats['iterations']} steps)", RuntimeWarning, stacklevel=3)
    torch.cuda.synchronize()
    stats = {"method": "galerkin", **stats, "wall_time": time.perf_counter() - t0}
    return ctx.dofmap.lift(u), stats
