"""Pins the oracle's reverse pass of the Galerkin solve (oracle/adjoint.py)
by central finite differences of the oracle's own direct solve: the CSR
reverse rule (the reference's _csr_sparse_solve_bwd,
src/jaxiga/solvers/linear.py:376-386) against perturbed values and load, and
the point cotangents against perturbed coefficients, sources, material
parameters and tractions -- with prescribed values that are not zero, so
that the lifting term is part of what is checked."""

import numpy as np
import scipy.sparse.linalg as spla

from oracle import adjoint as AJ
from oracle import assembly as A
from oracle import forms as F
from oracle import geometry as G
from oracle import pipeline as P
from oracle import space as S
from tests.cases import UNIT, everywhere

RTOL = 1e-6


def _fd(f, h):
    return (f(h) - f(-h)) / (2 * h)


def _poisson(V, a0_vals, amp, g=lambda x, p=None: 0.3 + x[0] - 0.5 * x[1]):
    return P.solve("poisson", V, [F.DirichletBC(g, everywhere)],
                   source=lambda x, p: p["amp"] * (1.0 + x[0] * x[1]),
                   params={"a0": F.PointField(a0_vals), "amp": amp}, method="direct")


def test_csr_solve_vjp_matches_finite_differences():
    rng = np.random.default_rng(0)
    V = S.function_space(G.quadrilateral(UNIT).elevate(2).refine(2))
    nq = 9
    out = _poisson(V, 1.0 + rng.random((V.n_elems, nq)), 2.0)
    pat, data, b = out["pattern"], out["data"], out["b"]
    c = rng.standard_normal(len(b))
    data_bar, b_bar = AJ.csr_solve_vjp(pat, data, out["u_free"], c)

    def loss(d, bb):
        return c @ spla.spsolve(A.to_scipy(pat, d).tocsc(), bb)

    dd = rng.standard_normal(pat.nnz)
    h = 1e-6 * np.abs(data).max()
    fd = _fd(lambda e: loss(data + e * dd, b), h)
    np.testing.assert_allclose(data_bar @ dd, fd, rtol=RTOL)
    db = rng.standard_normal(len(b))
    fd = _fd(lambda e: loss(data, b + e * db), 1e-6)
    np.testing.assert_allclose(b_bar @ db, fd, rtol=RTOL)


def test_poisson_point_cotangents_match_finite_differences():
    rng = np.random.default_rng(1)
    V = S.function_space(G.quadrilateral([[0, 0], [1.2, 0.1], [1.0, 1.0], [-0.1, 0.9]]).elevate(2).refine(2))
    nq = 9
    a0 = 1.0 + rng.random((V.n_elems, nq))
    amp = 1.5
    out = _poisson(V, a0, amp)
    c = rng.standard_normal(V.n_dofs)
    dm = out["dofmap"]
    data_bar, b_bar = AJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c[dm.free])
    gd = F.gdofs(V)
    Kbar, load_bar = AJ.element_cotangent(out["pattern"], data_bar, b_bar, gd, dm)
    ct = AJ.point_cotangents("poisson", out["basis"], Kbar, load_bar, gd, 1)

    da = rng.standard_normal(a0.shape)
    fd = _fd(lambda e: c @ _poisson(V, a0 + e * da, amp)["u"], 1e-6)
    np.testing.assert_allclose((ct["a0"] * da).sum(), fd, rtol=RTOL)
    x = out["basis"].x
    f_unit = 1.0 + x[..., 0] * x[..., 1]
    fd = _fd(lambda e: c @ _poisson(V, a0, amp + e)["u"], 1e-5)
    np.testing.assert_allclose((ct["source"][..., 0] * f_unit).sum(), fd, rtol=RTOL)


def _elastic(V, E, nu, t, plane="strain"):
    return P.solve("elasticity", V, [F.DirichletBC(lambda x, p=None: 0.01 * x[1], "left", component=0),
                                     F.DirichletBC(0.0, "left", component=1)],
                   [F.Neumann(lambda x, n, p: np.stack([p["t"] * (1.0 + 0 * x[0]), 0.3 * p["t"] * x[1]]),
                              "right")],
                   params={"E": E, "nu": nu, "t": t}, plane=plane, method="direct")


def test_elasticity_cotangents_match_finite_differences():
    rng = np.random.default_rng(2)
    V = S.function_space(G.rectangle(0, 0, 2, 1).elevate(2).refine(2), vec=2)
    E, nu, t = 1e3, 0.3, 2.0
    out = _elastic(V, E, nu, t)
    c = rng.standard_normal(V.n_dofs)
    dm = out["dofmap"]
    data_bar, b_bar = AJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c[dm.free])
    gd = F.gdofs(V)
    Kbar, load_bar = AJ.element_cotangent(out["pattern"], data_bar, b_bar, gd, dm)
    ct = AJ.point_cotangents("elasticity", out["basis"], Kbar, load_bar, gd, 2)
    Cf = lambda e_, n_: F.constitutive(e_, n_, 2, "strain")  # noqa: E731

    dE = (ct["C"] * Cf(1.0, nu)).sum()
    np.testing.assert_allclose(dE, _fd(lambda e: c @ _elastic(V, E + e, nu, t)["u"], 1e-3), rtol=RTOL)
    dC = _fd(lambda e: Cf(E, nu + e), 1e-6)
    np.testing.assert_allclose((ct["C"] * dC).sum(),
                               _fd(lambda e: c @ _elastic(V, E, nu + e, t)["u"], 1e-6), rtol=RTOL)
    bset = F.resolve_neumann(V, "right")
    bb = S.boundary_basis(V, bset)
    tb = AJ.boundary_cotangent(bb, load_bar, F.gdofs(V, bset.elems), 2)
    dt = np.stack([np.ones(bb.x.shape[:2]), 0.3 * bb.x[..., 1]], -1)
    np.testing.assert_allclose((tb * dt).sum(), _fd(lambda e: c @ _elastic(V, E, nu, t + e)["u"], 1e-5),
                               rtol=RTOL)


def _shape_case(kind, X):
    """An oracle solve on control points X: distorted Poisson quad (p=2, 4x4,
    a0 and source callables of x, constant non-zero Dirichlet value) or
    plane-strain elasticity on the rational quarter annulus (p=2, clamped
    inner edge, body force callable of x)."""
    if kind == "poisson":
        base = G.quadrilateral([[0, 0], [1.2, 0.1], [1.0, 1.0], [-0.1, 0.9]]).elevate(2).refine(2)
        V = S.function_space(G.Patch(X, base.weights, base.knots, base.degree, base.labels))
        params = {"a0": lambda x, p: 1.0 + x[0] ** 2 + 0.5 * x[1]}
        out = P.solve("poisson", V, [F.DirichletBC(0.3, everywhere)],
                      source=lambda x, p: 1.0 + x[0] * x[1], params=params, method="direct")
        return V, out, params, lambda x, p: 1.0 + x[0] * x[1]
    base = G.quarter_annulus(1.0, 2.0).elevate(2).refine(1)
    V = S.function_space(G.Patch(X, base.weights, base.knots, base.degree, base.labels), vec=2)
    params = {"E": 1.0, "nu": 0.3}
    src = lambda x, p: np.stack([0.2 + x[1], -1.0 + 0.3 * x[0] * x[1]])  # noqa: E731
    out = P.solve("elasticity", V, [F.DirichletBC(0.0, "inner")], source=src, params=params,
                  plane="strain", method="direct")
    return V, out, params, src


def _base_cpts(kind):
    if kind == "poisson":
        return G.quadrilateral([[0, 0], [1.2, 0.1], [1.0, 1.0], [-0.1, 0.9]]).elevate(2).refine(2).ctrl_pts
    return G.quarter_annulus(1.0, 2.0).elevate(2).refine(1).ctrl_pts


import pytest  # noqa: E402


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_shape_cotangent_matches_finite_differences(kind):
    """d(c.u)/dX along random directions of the control points (interior and
    boundary points both move) against central differences of the oracle's
    direct solves on the moved patches."""
    X0 = _base_cpts(kind)
    V, out, params, src = _shape_case(kind, X0)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(V.n_dofs)
    K = A.to_scipy(out["pattern"], out["data"])
    lam_free = spla.spsolve(K.tocsc(), c[out["dofmap"].free])
    lam = np.zeros(V.n_dofs)
    lam[out["dofmap"].free] = lam_free
    basis = S.gauss_basis(V)
    Gc = AJ.shape_cotangent(kind, V, basis, out["u"], lam, params, src,
                            plane="strain" if kind == "elasticity" else "stress")
    for trial in range(3):
        dX = rng.standard_normal(X0.shape)
        h = 1e-6
        fd = (c @ _shape_case(kind, X0 + h * dX)[1]["u"] - c @ _shape_case(kind, X0 - h * dX)[1]["u"]) / (2 * h)
        np.testing.assert_allclose((Gc * dX).sum(), fd, rtol=1e-5)
