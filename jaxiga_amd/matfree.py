"""Matrix-free operator and solve (sum-factorised, GPU): scalar Poisson and
linear elasticity.

Mirrors the matrix-free branch of the reference: ``tangent_matvec``
(``methods/galerkin.py:207-218``), ``tangent_diagonal`` (:221-237),
``matrix_free_solve`` (``solvers/linear.py:414-442``) and the
``matrix_free=True`` branch of ``solve_galerkin`` (:305-336). The reference
applies the operator as a JVP of the autodiff residual; here it is a
sum-factorised apply over per-Gauss-point geometric factors computed once in
FP64 (``iga_mf_setup``), with the apply arithmetic in FP64 or FP32 and the
Krylov recurrences in FP64. No CSR pattern is built on this path
(``stats["nnz"] is None``).
"""

from __future__ import annotations

import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd.runtime import lib, stream


def separable_map(space, rtol=1e-14):
    """True when the patch map is a tensor product of 1-D maps: control-point
    coordinate ``c`` depends only on the ``c``-th index (every refined box),
    to a relative spread of ``rtol`` (knot insertion leaves rounding-level
    differences)."""
    dim = space.dim
    if space.dim_phys != dim:
        return False
    n = tuple(space.patch.n_cp_per_dir)
    X = np.asarray(space.cpts, dtype=np.float64).reshape(*n[::-1], dim)
    scale = max(float(np.abs(X).max()), 1e-300)
    for c in range(dim):
        # index c is axis dim-1-c of the (i_{D-1}, .., i_0) layout
        Xc = np.moveaxis(X[..., c], dim - 1 - c, -1).reshape(-1, n[c])
        if np.abs(Xc - Xc[:1]).max() > rtol * scale:
            return False
    return True


class MatrixFreeOperator:
    """``y = K x`` on the free dofs of a scalar Poisson or a linear elasticity
    problem.

    ``precision``: ``IGA_FP64`` or ``IGA_FP32`` arithmetic, or ``IGA_BF16``
    (coefficient stored in BF16, FP32 arithmetic; separable patches only).
    The geometry is always built in FP64. On a separable polynomial patch
    (:func:`separable_map`) only the coefficient is stored per Gauss point
    and the Jacobi diagonal is sum-factorised (:meth:`diagonal`)."""

    def __init__(self, ctx, params=None, precision=_lib.IGA_FP64, inputs=None):
        """``inputs`` (optional): the values of :func:`jaxiga_amd.methods.point_inputs`;
        its coefficient is used (detached) instead of evaluating ``params``."""
        import torch

        from jaxiga_amd.methods import _bind_coefficient, _coefficient_args

        space = ctx.space
        kernel = ctx.problem.kernel
        self.elastic = kernel == "elasticity"
        if not ((kernel == "poisson" and space.vec == 1) or (self.elastic and space.vec == space.dim)):
            raise NotImplementedError("the matrix-free operator covers the scalar Poisson form and "
                                      "linear elasticity")
        if ctx.n_q1 != space.degree[0] + 1:
            raise NotImplementedError("the matrix-free operator uses the p+1 Gauss rule")
        self.ctx = ctx
        self.precision = precision
        # constant weights (any value) give R_i = N_i: the polynomial apply,
        # which skips the weight data and the quotient rule. Knot insertion
        # leaves weights of 1 a few ulps off, hence the relative tolerance;
        # the geometry factors always use the patch's own weights.
        w = np.asarray(space.wgts, dtype=np.float64)
        self.rational = int(w.size > 0 and np.ptp(w) > 1e-14 * np.abs(w).max())
        # elasticity uses the general layout (per-point inverse Jacobians)
        self.separable = int(not self.rational and not self.elastic and separable_map(space))
        self.flags = (_lib.IGA_MF_RATIONAL if self.rational else 0) | \
            (_lib.IGA_MF_SEPARABLE if self.separable else 0) | \
            (_lib.IGA_MF_ELASTICITY if self.elastic else 0)
        if precision == _lib.IGA_BF16 and not self.separable:
            raise NotImplementedError("the BF16 matrix-free operator needs a separable polynomial patch "
                                      "(scalar Poisson)")
        L = lib()
        st = space.device().struct
        nbytes = int(L.iga_mf_workspace(st, precision, self.flags))
        if nbytes < 0:
            raise NotImplementedError(f"no matrix-free kernel for dim={space.dim}, degree={space.degree}")
        self.work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=ctx.f2f_d.device)
        self.nbytes = nbytes
        if self.elastic:
            C = inputs["C"] if inputs is not None else ctx.problem.C(params)
            C = C.detach().cpu().numpy() if isinstance(C, torch.Tensor) else np.asarray(C, float)
            C = np.ascontiguousarray(C, dtype=np.float64)
            _lib.check(L.iga_mf_setup_elasticity(st, ctx.rule(), precision, self.flags, C.ctypes.data,
                                                 _lib.ptr(self.work), nbytes, stream()),
                       "iga_mf_setup_elasticity")
        else:
            args = _lib.IgaAsmArgs()
            keep = (_bind_coefficient(args, inputs["coef"]) if inputs is not None
                    else _coefficient_args(ctx, params, args))
            _lib.check(L.iga_mf_setup(st, ctx.rule(), precision, self.flags, args.coef_kind,
                                      args.coef_const, args.coef_points, _lib.ptr(self.work), nbytes,
                                      stream()), "iga_mf_setup")
            del keep
        self.n = ctx.n_free

    def apply(self, x, out=None):
        import torch

        y = torch.empty_like(x) if out is None else out
        ctx = self.ctx
        _lib.check(lib().iga_mf_apply(ctx.space.device().struct, _lib.ptr(self.work), self.nbytes,
                                      self.n, _lib.ptr(ctx.free_d), _lib.ptr(ctx.f2f_d), _lib.ptr(x),
                                      _lib.ptr(y), stream()), "iga_mf_apply")
        return y

    __call__ = apply

    def diagonal(self):
        """``diag(K)`` on the free dofs (FP64), sum-factorised; separable
        patches only (elsewhere the element pass gives it:
        :func:`jaxiga_amd.methods.tangent_diagonal`)."""
        import torch

        if not self.separable:
            raise NotImplementedError("the sum-factorised diagonal needs a separable polynomial patch")
        ctx = self.ctx
        d = torch.empty(self.n, dtype=torch.float64, device=ctx.f2f_d.device)
        _lib.check(lib().iga_mf_diag(ctx.space.device().struct, _lib.ptr(self.work), self.nbytes,
                                     self.n, _lib.ptr(ctx.free_d), _lib.ptr(ctx.f2f_d), _lib.ptr(d),
                                     stream()), "iga_mf_diag")
        return d

    def cg(self, b, tol=1e-10, maxiter=None, dinv=None, workspace=None):
        import torch

        L = lib()
        n = self.n
        if dinv is None:
            dinv = torch.ones(n, dtype=torch.float64, device=b.device)
        if workspace is None:
            workspace = torch.empty(int(L.iga_cg_workspace(n)), dtype=torch.float64, device=b.device)
        x = torch.empty(n, dtype=torch.float64, device=b.device)
        it, rr, bb = _lib.c_int(0), _lib.c_double(0.0), _lib.c_double(0.0)
        maxiter = 10 * n if maxiter is None else int(maxiter)
        ctx = self.ctx
        _lib.check(L.iga_cg_solve_mf(ctx.space.device().struct, _lib.ptr(self.work), self.nbytes, n,
                                     _lib.ptr(ctx.free_d), _lib.ptr(ctx.f2f_d), _lib.ptr(dinv),
                                     _lib.ptr(b), _lib.ptr(x), float(tol), maxiter, _lib.ptr(workspace),
                                     _lib.ctypes.byref(it), _lib.ctypes.byref(rr), _lib.ctypes.byref(bb),
                                     stream()), "iga_cg_solve_mf")
        return LIN.CGResult(x, it.value, rr.value, bb.value)


def inverse_guard(d):
    import torch

    inv = torch.empty_like(d)
    _lib.check(lib().iga_inv_guard(d.shape[0], _lib.ptr(d), _lib.ptr(inv), stream()), "iga_inv_guard")
    return inv


def diff_norms(y, b):
    """``(||y - b||, ||b||)`` in FP64 on the device."""
    import torch

    out = torch.zeros(2, dtype=torch.float64, device=b.device)
    work = torch.empty(2048, dtype=torch.float64, device=b.device)
    _lib.check(lib().iga_diff_norms(b.shape[0], _lib.ptr(y), _lib.ptr(b), _lib.ptr(out),
                                    _lib.ptr(work), stream()), "iga_diff_norms")
    o = out.cpu().numpy()
    return float(np.sqrt(o[0])), float(np.sqrt(o[1]))


def tangent_matvec(ctx, params=None, precision=_lib.IGA_FP64):
    return MatrixFreeOperator(ctx, params, precision)


def matrix_free_solve_ctx(ctx, params, linear, precision=_lib.IGA_FP64):
    """The ``matrix_free=True`` branch of the Galerkin driver."""
    import torch

    from jaxiga_amd.methods import assemble

    if linear.method != "cg":
        raise NotImplementedError(f"matrix-free method {linear.method!r} is not provided; use 'cg'")
    t0 = time.perf_counter()
    jacobi = linear.preconditioner == "jacobi"
    op = MatrixFreeOperator(ctx, params, precision)
    # the Jacobi diagonal: sum-factorised on separable patches, else from the
    # element pass that also forms the load
    _, b, diag = assemble(ctx, params, values=False, load=True,
                          diagonal=jacobi and not op.separable)
    if jacobi and op.separable:
        diag = op.diagonal()
    dinv = inverse_guard(diag) if diag is not None else None
    LIN.SOLVE_COUNTER["forward"] += 1
    res = op.cg(b, linear.tol, linear.maxiter, dinv=dinv)
    rn, bn = diff_norms(op.apply(res.x), b)
    torch.cuda.synchronize()
    stats = {"method": "galerkin", "linear": linear.method, "matrix_free": True,
             "n_free": ctx.n_free, "nnz": None, "iterations": res.iterations,
             "wall_time": time.perf_counter() - t0}
    stats.update(LIN.linear_diagnostics(rn, bn, res.x, linear.tol))
    LIN.warn_if_failed("linear Galerkin solve", stats["converged"], stats["residual"])
    return ctx.dofmap.lift(res.x), ctx, stats
