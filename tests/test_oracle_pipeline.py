"""Pins the oracle's end-to-end solve with the reference's known-answer tests:
patch test and constant Dirichlet data (tests/test_galerkin.py:35-53), the
configuration-1 problem (32x32 bicubic, 1225 dofs, 47961 nonzeros), and the
quarter annulus under internal pressure vs the Lame solution
(tests/test_examples.py:60-66)."""

import numpy as np

from oracle import forms as F
from oracle import geometry as G
from oracle import pipeline as P
from oracle import solver as SV
from oracle import space as S
from tests.cases import UNIT, everywhere, sin_exact, sin_source


def test_patch_test_and_constant_dirichlet():
    lin = lambda x, p=None: 1.0 + 2.0 * x[0] - 3.0 * x[1]  # noqa: E731
    for deg in (2, 3):
        V = S.function_space(G.quadrilateral(UNIT).elevate(deg).refine(2))
        out = P.solve("poisson", V, [F.DirichletBC(lin, everywhere)], source=lambda x, p: 0.0 * x[0],
                      params={"a0": 1.0}, method="direct")
        assert SV.errornorm(V, out["u"], lambda x: lin(x)) < 1e-12
    V = S.function_space(G.quadrilateral(UNIT).elevate(3).refine(2))
    out = P.solve("poisson", V, [F.DirichletBC(2.5, everywhere)], source=lambda x, p: 0.0 * x[0],
                  params={"a0": 1.0}, method="direct")
    np.testing.assert_allclose(out["u"], 2.5, atol=1e-11)


def test_configuration_1_sizes_and_accuracy():
    V = S.function_space(G.quadrilateral(UNIT).elevate(3).refine(5))
    out = P.solve("poisson", V, [F.DirichletBC(0.0, everywhere)], source=sin_source,
                  params={"a0": 1.0}, tol=1e-10)
    assert V.n_dofs == 1225 and out["stats"]["n_free"] == 1089 and out["stats"]["nnz"] == 47961
    assert out["stats"]["converged"]
    e = SV.errornorm(V, out["u"], sin_exact)
    # p=3, h=1/32: O(h^4) L2 error well below 1e-5
    assert 1e-7 < e < 1e-5


def test_lame_annulus():
    a2, b2, pr, E, nu = 1.0, 16.0, 10.0, 1e5, 0.3

    def disp(x):
        r = np.hypot(x[0], x[1])
        th = np.arctan2(x[1], x[0])
        ur = pr * a2 * r / (E * (b2 - a2)) * ((1 - nu) + (1 + nu) * b2 / r**2)
        return np.stack([ur * np.cos(th), ur * np.sin(th)])

    V = S.function_space(G.quarter_annulus(1.0, 4.0).elevate(4).refine(3), vec=2)
    out = P.solve("elasticity", V,
                  [F.DirichletBC(0.0, "sym_y", component=1), F.DirichletBC(0.0, "sym_x", component=0)],
                  [F.Neumann(lambda x, n, p: -pr * n, "inner")], params={"E": E, "nu": nu},
                  plane="stress", method="direct")
    assert SV.errornorm(V, out["u"], disp) < 5e-5
