"""The allocation-free CG entry point (iga_cg_solve_noalloc, no reference
counterpart) against iga_cg_solve on the same system: two calls in a row,
the second on a smaller system, each with the same iteration count and a
bitwise-equal solution (same kernels, fixed reduction order); sizes above
the stated maximum are rejected."""

import ctypes

import numpy as np
import pytest
import torch

from jaxiga_amd import _lib
from jaxiga_amd import grf
from jaxiga_amd import linear as LIN
from jaxiga_amd.runtime import lib, stream

pytestmark = pytest.mark.gpu


def _system(refine, seed):
    ctx, vals, rhs = grf.poisson_instance(refine, seed, degree=2)
    return ctx.assembly, vals, rhs


def test_noalloc_cg_matches_cg_solve(gpu):
    for refine, seed, prec in ((6, 1, torch.float64), (4, 2, torch.float64), (5, 3, torch.float32)):
        pat, vals, rhs = _system(refine, seed)
        vals = vals.to(prec)
        print(f"[noalloc] refine={refine} n={pat.shape[0]} dtype={prec}", flush=True)
        dinv = LIN.inverse_diagonal(pat, vals.double())
        ref = LIN.cg_csr(pat, vals, rhs, tol=1e-10, dinv=dinv)
        x = torch.full_like(rhs, float("nan"))
        got = LIN.cg_csr_noalloc(pat, vals, rhs, x, dinv, tol=1e-10)
        torch.cuda.synchronize()
        assert got.iterations == ref.iterations > 0
        assert torch.equal(got.x, ref.x)
        assert got.rr == ref.rr and got.bb == ref.bb
        rn, bn = LIN.residual_norms(pat, vals.double(), got.x, rhs)
        assert rn <= 1e-9 * bn


def test_noalloc_cg_rejects_sizes_beyond_its_maximum(gpu):
    n = LIN.NOALLOC_MAX_N + 1
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    it, rr, bb = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    rc = lib().iga_cg_solve_noalloc(n, _lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy),
                                    _lib.IGA_FP64, _lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy),
                                    1e-10, 10, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(bb),
                                    stream())
    assert rc == _lib.IGA_ERR_SHAPE
