"""Mixed precision with FP64 accuracy recovered on the FP64 operator.

The reference is FP64 only (``src/jaxiga/config.py:1-33``). This module adds
the low-precision operators of this build with FP64 accuracy recovered in
two ways.

Reliable updates (:func:`csr_mixed_cg`, the default for CSR): one Jacobi-PCG
run on the FP32 matrix with FP64 vectors and recurrences; every few
iterations, and whenever the recursion reports convergence, the residual is
replaced by the true ``b - K64 x``, keeping the search direction, so the
Krylov space is never discarded and the stopping test holds for K64.

Iterative refinement (:func:`csr_mixed_solve`, :func:`mf_mixed_solve`):

* outer loop in FP64: residual ``r = b - K64 u`` from the FP64 operator (CSR
  or matrix-free), convergence judged on ``||r|| <= tol ||b||``;
* inner solve ``K_lo d = r`` by Jacobi-PCG whose operator is the FP32 CSR
  values (FP32 matrix cores at assembly) or the FP32 matrix-free apply, with
  FP64 vectors, dot products and recurrences, to a loose relative tolerance;
* ``u += d``.

The FP64 steps are exactly the ones the design keeps in FP64: geometry
(Jacobian, determinant, weights), the outer residual and every Krylov
recurrence.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd.runtime import lib, stream


@dataclasses.dataclass
class RefinementResult:
    x: object
    outer: int
    inner_iterations: list
    residuals: list  # relative FP64 residual after each outer step
    seconds: float

    @property
    def total_inner(self):
        return int(sum(self.inner_iterations))


def _axpy(a, x, y):
    _lib.check(lib().iga_axpy(x.shape[0], float(a), _lib.ptr(x), _lib.ptr(y), stream()), "iga_axpy")


def csr_residual(pattern, vals64, u, b, r, work):
    """``r = b - K u`` and ``(||r||, ||b||)``; FP64."""
    import torch

    out = torch.empty(2, dtype=torch.float64, device=b.device)
    _lib.check(lib().iga_residual_vec(pattern.shape[0], _lib.ptr(pattern.indptr_d),
                                      _lib.ptr(pattern.cols_d), _lib.ptr(vals64), _lib.ptr(u),
                                      _lib.ptr(b), _lib.ptr(r), _lib.ptr(out), _lib.ptr(work),
                                      stream()), "iga_residual_vec")
    o = out.cpu().numpy()
    return float(np.sqrt(o[0])), float(np.sqrt(o[1]))


def refine(residual_fn, inner_fn, b, tol=1e-10, inner_tol=1e-6, max_outer=12):
    """Generic FP64 iterative refinement.

    ``residual_fn(u, r) -> (||r||, ||b||)`` writes ``r = b - K64 u``;
    ``inner_fn(r, rel_tol) -> (d, iterations)`` solves the low-precision
    correction system.
    """
    import torch

    t0 = time.perf_counter()
    u = torch.zeros_like(b)
    r = b.clone()
    bn = None
    rn, bn = residual_fn(u, r)
    inner, hist = [], [rn / (bn if bn > 0 else 1.0)]
    k = 0
    while hist[-1] > tol and k < max_outer:
        # ask the inner solve for what is left, but no tighter than its precision can give
        want = max(inner_tol, 0.5 * tol * bn / max(rn, 1e-300))
        d, its = inner_fn(r, min(want, 0.5))
        _axpy(1.0, d, u)
        rn, bn = residual_fn(u, r)
        inner.append(its)
        hist.append(rn / (bn if bn > 0 else 1.0))
        k += 1
    torch.cuda.synchronize()
    return RefinementResult(u, k, inner, hist, time.perf_counter() - t0)


@dataclasses.dataclass
class MixedCGResult:
    x: object
    iterations: int
    relative_residual: float  # true FP64 ||b - K64 x|| / ||b||
    seconds: float


def csr_mixed_cg(pattern, vals64, vals32, b, dinv, tol=1e-10, maxiter=None, replace_every=64,
                 workspace=None):
    """Jacobi-PCG on the FP32 CSR values with FP64 vectors and reliable
    updates against the FP64 values (same pattern); stops on the true FP64
    residual."""
    import torch

    L = lib()
    n = pattern.shape[0]
    if workspace is None:
        workspace = torch.empty(int(L.iga_cg_workspace(n)), dtype=torch.float64, device=b.device)
    x = torch.empty(n, dtype=torch.float64, device=b.device)
    it = _lib.c_int(0)
    rr = _lib.c_double(0.0)
    bb = _lib.c_double(0.0)
    maxiter = 10 * n if maxiter is None else int(maxiter)
    t0 = time.perf_counter()
    _lib.check(L.iga_cg_solve_mixed(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d),
                                    _lib.ptr(vals32), _lib.ptr(vals64), _lib.ptr(dinv), _lib.ptr(b),
                                    _lib.ptr(x), float(tol), maxiter, int(replace_every),
                                    _lib.ptr(workspace), _lib.ctypes.byref(it),
                                    _lib.ctypes.byref(rr), _lib.ctypes.byref(bb), stream()),
               "iga_cg_solve_mixed")
    rel = float(np.sqrt(rr.value / bb.value)) if bb.value > 0 else 0.0
    return MixedCGResult(x, it.value, rel, time.perf_counter() - t0)


def csr_mixed_solve(pattern, vals64, vals32, b, dinv, tol=1e-10, inner_tol=1e-6, maxiter=None):
    """FP32-CSR inner Jacobi-PCG, FP64-CSR outer residual."""
    import torch

    n = pattern.shape[0]
    work = torch.empty(2048, dtype=torch.float64, device=b.device)
    cgw = torch.empty(int(lib().iga_cg_workspace(n)), dtype=torch.float64, device=b.device)

    def residual_fn(u, r):
        return csr_residual(pattern, vals64, u, b, r, work)

    def inner_fn(r, rel):
        res = LIN.cg_csr(pattern, vals32, r, rel, maxiter, dinv=dinv, workspace=cgw)
        return res.x, res.iterations

    return refine(residual_fn, inner_fn, b, tol, inner_tol)


def mf_mixed_solve(op64, op32, b, dinv, tol=1e-10, inner_tol=1e-6, maxiter=None):
    """FP32 matrix-free inner Jacobi-PCG, FP64 matrix-free outer residual."""
    import torch

    from jaxiga_amd.matfree import diff_norms

    n = b.shape[0]
    cgw = torch.empty(int(lib().iga_cg_workspace(n)), dtype=torch.float64, device=b.device)
    Ku = torch.empty_like(b)

    def residual_fn(u, r):
        op64.apply(u, out=Ku)
        rn, bn = diff_norms(Ku, b)
        torch.sub(b, Ku, out=r)
        return rn, bn

    def inner_fn(r, rel):
        res = op32.cg(r, rel, maxiter, dinv=dinv, workspace=cgw)
        return res.x, res.iterations

    return refine(residual_fn, inner_fn, b, tol, inner_tol)
