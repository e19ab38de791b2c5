"""Mixed precision with FP64 iterative refinement.

The reference is FP64 only (``src/jaxiga/config.py:1-33``). This module adds
the low-precision operators of this build with FP64 accuracy recovered by
iterative refinement:

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
