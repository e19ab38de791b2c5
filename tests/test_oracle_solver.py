"""Pins the oracle's CG variants, diagnostics and norms (reference
tests/test_galerkin.py:71-79,399-406, tests/test_solver_robustness.py:85-93,
tests/test_advanced.py:240-269)."""

import numpy as np
import scipy.sparse.linalg as spla

from oracle import geometry as G
from oracle import pipeline as P
from oracle import solver as SV
from oracle import space as S
from tests.cases import everywhere, sin_exact, sin_source
from oracle import forms as F


def _cfg(ref=3):
    V = S.function_space(G.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(3).refine(ref))
    return V, P.solve("poisson", V, [F.DirichletBC(0.0, everywhere)], source=sin_source,
                      params={"a0": 1.0}, method="direct")


def test_cg_variants_agree_with_direct_solve():
    V, out = _cfg()
    from oracle import assembly as A

    K = A.to_scipy(out["pattern"], out["data"])
    b = out["b"]
    ref = spla.spsolve(K.tocsc(), b)
    u1, it1 = SV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-13)
    u2, it2 = SV.scipy_csr_cg(K, b, tol=1e-13)
    np.testing.assert_allclose(u1, ref, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(u2, ref, rtol=1e-7, atol=1e-9)
    assert it1 > 0 and it2 > 0
    assert np.linalg.norm(K @ u2 - b) <= 1e-13 * np.linalg.norm(b)


def test_truncated_cg_is_flagged_not_raised():
    V, out = _cfg()
    from oracle import assembly as A

    K = A.to_scipy(out["pattern"], out["data"])
    u, it = SV.jacobi_pcg(lambda v: K @ v, out["b"], K.diagonal(), tol=1e-10, maxiter=1)
    assert it == 1
    d = SV.linear_diagnostics(lambda v: K @ v, u, out["b"], 1e-10)
    assert not d["converged"] and d["finite"]


def test_norm_of_sin_sin_is_one_half():
    V, out = _cfg(3)
    np.testing.assert_allclose(SV.l2_norm(V, out["u"]), 0.5, rtol=1e-3)
    np.testing.assert_allclose(SV.integrate(V, lambda x: np.ones(x.shape[1])), 1.0, rtol=1e-12)


def test_errornorm_convergence_rate():
    errs = []
    for r in (2, 3, 4):
        V, out = _cfg(r)
        errs.append(SV.errornorm(V, out["u"], sin_exact))
    assert errs[0] > errs[1] > errs[2]
    assert np.log2(errs[1] / errs[2]) > 3.5
