"""Implicit differentiation of the Galerkin solve (``jax.grad`` through
``jx.solve`` in the reference).

The reference differentiates ``solve`` by JAX autodiff of the assembly plus
the custom reverse rule of the sparse solve (``_csr_sparse_solve``,
``src/jaxiga/solvers/linear.py:362-389``): one adjoint solve ``K lam = u_bar``
on the same symmetric CSR, ``data_bar = -lam[rows] u[cols]`` and
``b_bar = lam``, which then flow back through the scatter-add and the element
kernels (``src/jaxiga/methods/galerkin.py:37-148``) to the parameters.

Here the parameters reach the element kernels only as values at points: the
coefficient ``a0`` (a number or its values at the interior Gauss points), the
constitutive matrix ``C``, the source at the interior Gauss points and the
tractions at the boundary Gauss points (:func:`jaxiga_amd.methods.point_inputs`,
evaluated with torch autograd on). Because ``data_bar`` is the rank-one
``-lam u^T`` on the pattern and the load enters linearly, its pull-back through
the scatter-add and the element kernels is a sum over points of the two
fields, with ``u`` the full solution (prescribed values included, which
carries the lifting ``-K_fp g``) and ``lam`` zero on prescribed dofs:

* coefficient at a point: ``-w grad(lam) . grad(u)``;
* ``C``: ``-sum_points w eps(lam) eps(u)^T`` (Voigt, engineering shear);
* source at a point: ``w lam``; traction at a boundary point: ``w lam``.

``lam`` and ``u`` at the points come from one HIP kernel
(``iga_field_points``) over the basis data of the same points. The reverse
pass therefore costs exactly one adjoint solve (``SOLVE_COUNTER['adjoint']``),
whatever the forward iteration did, and torch carries the point cotangents
back to the parameters through the user's callables. Dirichlet values are
resolved at set-up and are not differentiated (as in the reference).

Control points (a space from ``FunctionSpace.replace(cpts=tensor)``) get the
shape cotangent :meth:`_Run.shape_cotangent` from the same adjoint field:
the first variation of the point functional under a motion of the map,
pulled back to the control points by ``iga_point_forces``. Weights are not
differentiated. With
``matrix_free=True`` both solves are the matrix-free CG (the symmetric
operator is its own transpose, ``src/jaxiga/solvers/linear.py:436-442``).
"""

from __future__ import annotations

import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd.evaluation import evaluate
from jaxiga_amd.pointset import gauss
from jaxiga_amd.runtime import lib, stream


def field_points(basis, u0, u1=None, vec=1, values=True, gradients=True):
    """Values ``(n_f, n_e, n_q, vec)`` and gradients ``(n_f, n_e, n_q, vec,
    dim)`` at the points of ``basis`` of one or two full coefficient vectors
    (``n_cp * vec`` doubles, component fastest); either output may be
    skipped (None)."""
    import torch

    if values and basis.R is None:
        raise ValueError("values need the basis functions R")
    if gradients and basis.dR is None:
        raise ValueError("gradients need the basis gradients dR")
    ref = basis.R if basis.R is not None else basis.dR
    n_e, n_q, n_l = ref.shape[0], ref.shape[1], ref.shape[-1]
    dim = basis.dR.shape[2] if basis.dR is not None else _dim(basis)
    dev = u0.device
    nf = 1 if u1 is None else 2
    dofs = basis.__dict__.get("_dofs_dev")
    if dofs is None or dofs.device != dev:
        # the element dof table on the device, kept with the basis data
        dofs = torch.as_tensor(np.ascontiguousarray(basis.dofs, dtype=np.int32), device=dev)
        basis.__dict__["_dofs_dev"] = dofs
    u0 = u0.detach().to(torch.float64).contiguous()
    u1 = None if u1 is None else u1.detach().to(torch.float64).contiguous()
    val = torch.empty((nf, n_e, n_q, vec), dtype=torch.float64, device=dev) if values else None
    grad = torch.empty((nf, n_e, n_q, vec, dim), dtype=torch.float64, device=dev) if gradients else None
    _lib.check(lib().iga_field_points(n_e, n_q, n_l, dim, vec, _lib.ptr(basis.R) if values else None,
                                      _lib.ptr(basis.dR) if gradients else None, _lib.ptr(dofs),
                                      _lib.ptr(u0), _lib.ptr(u1), _lib.ptr(val), _lib.ptr(grad),
                                      stream()), "iga_field_points")
    return val, grad


def _dim(basis):
    for a in (basis.x, basis.normal):
        if a is not None:
            return int(a.shape[-1])
    raise ValueError("the basis data carries no spatial dimension (evaluate 'dR' or 'x')")


def voigt(g):
    """Engineering Voigt strain ``(..., n_s)`` from displacement gradients
    ``g[..., c, d] = du_c/dx_d`` (2-D: xx, yy, xy; 3-D: xx, yy, zz, yz, xz, xy)."""
    import torch

    if g.shape[-1] == 2:
        return torch.stack([g[..., 0, 0], g[..., 1, 1], g[..., 0, 1] + g[..., 1, 0]], -1)
    return torch.stack([g[..., 0, 0], g[..., 1, 1], g[..., 2, 2], g[..., 1, 2] + g[..., 2, 1],
                        g[..., 0, 2] + g[..., 2, 0], g[..., 0, 1] + g[..., 1, 0]], -1)


class _Run:
    """One differentiable Galerkin solve: the forward assembly and solve, and
    the reverse pass (one adjoint solve + the point cotangents)."""

    def __init__(self, ctx, params, linear, inputs):
        self.ctx, self.params, self.linear, self.inputs = ctx, params, linear, inputs
        self.stats = None

    def forward(self):
        import torch

        from jaxiga_amd.methods import assemble

        ctx, linear = self.ctx, self.linear
        t0 = time.perf_counter()
        with torch.no_grad():
            if linear.matrix_free:
                u_free, info, rn, bn = self._forward_matrix_free()
            else:
                vals, b, _ = assemble(ctx, self.params, inputs=self.inputs)
                if linear.preconditioner == "multigrid":
                    from jaxiga_amd import learned as LN

                    u_free, info = LN.linear_solve(ctx, vals, b, linear)
                    self._adjoint = lambda ub: LN.linear_solve(ctx, vals, ub, linear, adjoint=True)[0]
                else:
                    u_free, info = LIN.sparse_solve(linear, ctx.assembly, vals, b)
                    # the reverse pass: one solve on the same symmetric CSR
                    self._adjoint = lambda ub: LIN.adjoint_solve(linear, ctx.assembly, vals, ub)
                rn, bn = LIN.residual_norms(ctx.assembly, vals, u_free, b)
            u = ctx.dofmap.lift(u_free)
        torch.cuda.synchronize()
        stats = {"method": "galerkin", "linear": linear.method, "matrix_free": bool(linear.matrix_free),
                 "n_free": ctx.n_free, "nnz": None if linear.matrix_free else ctx.assembly.nnz, **info,
                 "wall_time": time.perf_counter() - t0}
        stats.update(LIN.linear_diagnostics(rn, bn, u_free, linear.tol))
        LIN.warn_if_failed("linear Galerkin solve", stats["converged"], stats["residual"])
        self.stats, self.u = stats, u
        return u

    def _forward_matrix_free(self):
        """The matrix-free branch (``src/jaxiga/solvers/linear.py:414-442``):
        the adjoint of the symmetric operator is the same matrix-free CG."""
        from jaxiga_amd.matfree import MatrixFreeOperator, diff_norms, inverse_guard
        from jaxiga_amd.methods import assemble

        ctx, linear = self.ctx, self.linear
        if linear.method != "cg":
            raise NotImplementedError(f"matrix-free method {linear.method!r} is not provided; use 'cg'")
        _, b, diag = assemble(ctx, self.params, values=False, load=True,
                              diagonal=linear.preconditioner == "jacobi", inputs=self.inputs)
        op = MatrixFreeOperator(ctx, self.params, inputs=self.inputs)
        dinv = inverse_guard(diag) if diag is not None else None
        LIN.SOLVE_COUNTER["forward"] += 1
        res = op.cg(b, linear.tol, linear.maxiter, dinv=dinv)
        rn, bn = diff_norms(op.apply(res.x), b)

        def adjoint(ub):
            LIN.SOLVE_COUNTER["adjoint"] += 1
            return op.cg(ub.contiguous(), linear.tol, linear.maxiter, dinv=dinv).x

        self._adjoint = adjoint
        return res.x, {"iterations": res.iterations}, rn, bn

    def backward(self, u_bar, needs, shape=False):
        """Cotangents of ``(coef, src, C, *tractions)`` (None where not needed),
        and of the control points (``shape``) as an extra last item."""
        import torch

        ctx = self.ctx
        space = ctx.space
        dev = u_bar.device
        free = getattr(ctx, "_free_long", None)
        if free is None:
            free = ctx.free_d.long()
            ctx._free_long = free
        lam_free = self._adjoint(u_bar[free])
        lam = torch.zeros(space.n_dofs, dtype=torch.float64, device=dev)
        lam[free] = lam_free
        coef, src, C = self.inputs["coef"], self.inputs["src"], self.inputs["C"]
        tractions = self.inputs["tractions"] or []
        out = [None] * (3 + len(tractions))
        need_grad = needs[0] or needs[2]
        if needs[0] or needs[1] or needs[2]:
            outputs = ("R",) * bool(needs[1]) + ("dR",) * bool(need_grad) + ("x", "w")
            basis = evaluate(space, gauss(space, ctx.n_q1), outputs=outputs)
            w = basis.w
            if need_grad:
                _, g = field_points(basis, self.u, lam, space.vec, values=False)
                if needs[0]:
                    gq = -w * (g[1] * g[0]).sum((-1, -2))
                    out[0] = gq if coef.dim() == 2 else gq.sum().reshape(coef.shape)
                if needs[2]:
                    eu, el = voigt(g[0]), voigt(g[1])
                    gc = -torch.einsum("eqi,eqj->ij", el * w[..., None], eu)
                    out[2] = gc.to(C.device).reshape(C.shape)
            if needs[1]:
                v, _ = field_points(basis, lam, vec=space.vec, gradients=False)
                out[1] = w[..., None] * v[0]
        for k, t in enumerate(tractions):
            if not needs[3 + k]:
                continue
            from jaxiga_amd.methods import _neumann_sets

            _, ps, _ = _neumann_sets(ctx, dev)[k]
            basis = evaluate(space, ps, outputs=("R", "x", "w"))
            v, _ = field_points(basis, lam, vec=space.vec, gradients=False)
            out[3 + k] = basis.w[..., None] * v[0]
        if shape:
            out.append(self.shape_cotangent(lam))
        return out

    def shape_cotangent(self, lam):
        """``d/dX [lam^T (F - K u)]`` over the control points ``(n_cp, dim)``.
        Under a motion ``dX_k`` of control point k the Gauss points move by
        ``R_k dX_k``, so ``dw = w div V`` and ``d grad(phi) = -grad(V)^T
        grad(phi)``; the point functional ``w (f.lam - eps(lam):sigma(u))``
        then varies by ``P : grad(R_k) + R_k xbar``,
        ``P_cj = w [(f.lam - eps(lam):sigma(u)) delta_cj + sum_i (d_c lam_i
        sigma_ij(u) + d_c u_i sigma_ij(lam))]`` (``sigma = a grad u`` for
        Poisson), ``xbar`` the derivative through the explicit ``x`` of
        callable coefficients and sources (torch autograd of the callables).
        The pull-back to the control points is ``iga_point_forces``."""
        import torch

        ctx, space = self.ctx, self.ctx.space
        dim, vec = space.dim, space.vec
        basis = evaluate(space, gauss(space, ctx.n_q1), outputs=("R", "dR", "x", "w"))
        val, g = field_points(basis, self.u, lam, vec)
        w = basis.w
        gu, gl, lq = g[0], g[1], val[1]  # (n_e, n_q, vec, dim), (n_e, n_q, vec)
        src = self.inputs["src"]
        if src is None:
            fl = torch.zeros_like(w)
        else:
            f = torch.as_tensor(src, dtype=torch.float64, device=w.device).detach()
            fl = (f * lq).sum(-1)
        eye = torch.eye(dim, dtype=torch.float64, device=w.device)
        if ctx.problem.kernel == "poisson":
            coef = self.inputs["coef"]
            a = torch.ones_like(w) if coef is None else torch.as_tensor(coef, dtype=torch.float64,
                                                                          device=w.device).detach()
            a = a.expand_as(w) if a.dim() == 0 else a
            u1, l1 = gu[:, :, 0, :], gl[:, :, 0, :]
            inner = (u1 * l1).sum(-1)
            P = (fl - a * inner)[..., None, None] * eye
            P = P + a[..., None, None] * (l1[..., :, None] * u1[..., None, :] + u1[..., :, None] * l1[..., None, :])
        else:
            C = torch.as_tensor(self.inputs["C"], dtype=torch.float64).detach().to(w.device)

            def stress(gr):
                sv = voigt(gr) @ C.T
                if dim == 2:
                    rows = [[sv[..., 0], sv[..., 2]], [sv[..., 2], sv[..., 1]]]
                else:
                    rows = [[sv[..., 0], sv[..., 5], sv[..., 4]], [sv[..., 5], sv[..., 1], sv[..., 3]],
                            [sv[..., 4], sv[..., 3], sv[..., 2]]]
                return torch.stack([torch.stack(r, -1) for r in rows], -2)

            su, sl = stress(gu), stress(gl)
            energy = (voigt(gl) * (voigt(gu) @ C.T)).sum(-1)
            P = (fl - energy)[..., None, None] * eye
            P = P + torch.einsum("eqic,eqij->eqcj", gl, su) + torch.einsum("eqic,eqij->eqcj", gu, sl)
        P = (P * w[..., None, None]).contiguous()
        xbar = self._explicit_x(basis, lq, gu, gl)
        G = torch.zeros((space.n_scalar_basis, dim), dtype=torch.float64, device=w.device)
        dofs = basis.__dict__.get("_dofs_dev")
        _lib.check(lib().iga_point_forces(w.shape[0], w.shape[1], basis.dR.shape[-1], dim,
                                          _lib.ptr(basis.R), _lib.ptr(basis.dR), _lib.ptr(P),
                                          _lib.ptr(xbar), _lib.ptr(dofs), _lib.ptr(G), stream()),
                   "iga_point_forces")
        return G

    def _explicit_x(self, basis, lq, gu, gl):
        """``w d/dx (f(x).lam - a(x) grad(lam).grad(u))`` at the points through
        the callables' own dependence on x (None when nothing depends on x)."""
        import torch

        from jaxiga_amd.methods import call_points

        ctx, params = self.ctx, self.params
        problem = ctx.problem
        a0 = params.get("a0") if isinstance(params, dict) else None
        from jaxiga_amd.forms import PointField

        a_fn = a0 if (problem.kernel == "poisson" and callable(a0) and not isinstance(a0, PointField)) else None
        f_fn = problem.source if callable(problem.source) else None
        if a_fn is None and f_fn is None:
            return None
        x = basis.x.detach().clone().requires_grad_(True)
        with torch.enable_grad():
            S = torch.zeros((), dtype=torch.float64, device=x.device)
            if f_fn is not None:
                f = call_points(f_fn, x, params=params, vec=ctx.space.vec)
                S = S + (basis.w[..., None] * f * lq).sum()
            if a_fn is not None:
                a = call_points(a_fn, x, params=params)[..., 0]
                inner = (gu[:, :, 0, :] * gl[:, :, 0, :]).sum(-1)
                S = S - (basis.w * a * inner).sum()
            if not S.requires_grad:
                return None
            (xbar,) = torch.autograd.grad(S, x)
        return xbar.contiguous()


def _galerkin_fn():
    import torch

    class GalerkinSolve(torch.autograd.Function):
        @staticmethod
        def forward(fc, run, cpts, coef, src, C, *tractions):
            fc.run = run
            fc.cpts_meta = None if cpts is None else (cpts.shape, cpts.dtype, cpts.device)
            return run.forward()

        @staticmethod
        def backward(fc, u_bar):
            shape = bool(fc.needs_input_grad[1])
            grads = fc.run.backward(u_bar.detach().to(torch.float64).contiguous(), fc.needs_input_grad[2:],
                                    shape=shape)
            g_cpts = None
            if shape:
                shp, dt, dev = fc.cpts_meta
                g_cpts = grads.pop().to(device=dev, dtype=dt).reshape(shp)
            return (None, g_cpts, *grads)

    return GalerkinSolve


_FN = []


def galerkin_solve(ctx, params, linear, inputs):
    """``(u, stats)``: the full solution as a tensor that carries the
    gradient of every tensor in ``inputs`` (see the module docstring)."""
    import torch

    if not _FN:
        _FN.append(_galerkin_fn())
    run = _Run(ctx, params, linear, inputs)

    def as_arg(v):
        return v if isinstance(v, torch.Tensor) and v.requires_grad else None

    tractions = inputs["tractions"] or []
    u = _FN[0].apply(run, as_arg(getattr(ctx.space, "_cpts_t", None)), as_arg(inputs["coef"]),
                     as_arg(inputs["src"]), as_arg(inputs["C"]), *[as_arg(t) for t in tractions])
    return u, run.stats
