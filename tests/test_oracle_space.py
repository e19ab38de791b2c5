"""Pins the oracle's extraction, Bernstein, Gauss, DOF-map and evaluation
functions against SciPy, closed forms and exact integrals (the checks of the
reference's tests/test_bernstein.py, test_evaluation.py and
test_function_space.py:180-197)."""

import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import geometry as G
from oracle import space as S

DEGREES = [1, 2, 3, 4, 5]


def open_knots(deg, n_elem, repeat_interior=0):
    interior = np.repeat(np.linspace(0, 1, n_elem + 1)[1:-1], 1 + repeat_interior)
    return np.concatenate([np.zeros(deg + 1), interior, np.ones(deg + 1)])


@pytest.mark.parametrize("deg", DEGREES)
@pytest.mark.parametrize("rep", [0, 1])
def test_extraction_matches_scipy_design_matrix(deg, rep):
    kv = open_knots(deg, 5, rep if deg > 1 else 0)
    ops = S.bezier_extraction(kv, deg)
    lo, hi, first = S.element_spans(kv, deg)
    assert len(ops) == len(lo)
    u = np.linspace(-0.95, 0.95, 9)
    for e in range(len(lo)):
        x = lo[e] + (u + 1) / 2 * (hi[e] - lo[e])
        D = BSpline.design_matrix(x, kv, deg).toarray()[:, first[e] : first[e] + deg + 1]
        np.testing.assert_allclose(S.bernstein(u, deg) @ ops[e].T, D, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("deg", DEGREES)
def test_bernstein_matches_scipy(deg):
    u = np.linspace(-1, 1, 17)
    kv = np.concatenate([-np.ones(deg + 1), np.ones(deg + 1)])
    for k in range(deg + 1):
        c = np.zeros(deg + 1)
        c[k] = 1.0
        spl = BSpline(kv, c, deg)
        np.testing.assert_allclose(S.bernstein(u, deg)[:, k], spl(u), atol=1e-14)
        np.testing.assert_allclose(S.bernstein(u, deg, 1)[:, k], spl.derivative(1)(u), atol=1e-12)
        if deg >= 2:
            np.testing.assert_allclose(S.bernstein(u, deg, 2)[:, k], spl.derivative(2)(u), atol=1e-11)


@pytest.mark.parametrize("deg", DEGREES)
def test_bernstein_partition_of_unity_and_fd(deg):
    u = np.linspace(-0.8, 0.8, 11)
    np.testing.assert_allclose(S.bernstein(u, deg).sum(1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(S.bernstein(u, deg, 1).sum(1), 0.0, atol=1e-12)
    h = 1e-6
    fd = (S.bernstein(u + h, deg) - S.bernstein(u - h, deg)) / (2 * h)
    np.testing.assert_allclose(S.bernstein(u, deg, 1), fd, rtol=1e-6, atol=1e-8)


def test_tensor_numbering_first_direction_fastest():
    pts = np.array([[0.3, -0.2]])
    B, dB = S.bernstein_tensor(pts, (2, 1))
    bu, bv = S.bernstein(pts[:, 0], 2)[0], S.bernstein(pts[:, 1], 1)[0]
    np.testing.assert_allclose(B[0], np.kron(bv, bu), rtol=1e-15)
    du = S.bernstein(pts[:, 0], 2, 1)[0]
    np.testing.assert_allclose(dB[0, 0], np.kron(bv, du), rtol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gauss_rule_integrates_polynomials(n):
    pts, wts = S.gauss_rule([n, n])
    for a in range(2 * n):
        for b in range(2 * n):
            exact = (1 - (-1) ** (a + 1)) / (a + 1) * (1 - (-1) ** (b + 1)) / (b + 1)
            np.testing.assert_allclose((wts * pts[:, 0] ** a * pts[:, 1] ** b).sum(), exact, atol=1e-13)
    if n > 1:  # first direction fastest
        assert pts[1, 0] != pts[0, 0]
        assert pts[1, 1] == pts[0, 1]


@pytest.mark.parametrize("degree", [2, 3, 4])
@pytest.mark.parametrize("n_refine", [0, 1, 2])
def test_closed_form_ien_and_boundary_dofs(degree, n_refine):
    V = S.function_space(G.rectangle(0, 0, 2, 1).elevate(degree).refine(n_refine))
    spans = 2**n_refine
    n = spans + degree
    assert V.n_elems == spans**2 and V.n_scalar_basis == n**2
    idx = np.arange(n**2).reshape(n, n)
    for label, exp in (("left", idx[:, 0]), ("right", idx[:, -1]), ("bottom", idx[0]), ("top", idx[-1])):
        np.testing.assert_array_equal(np.sort(V.boundaries[label].dofs), np.sort(exp))
    for v in range(spans):
        for u in range(spans):
            np.testing.assert_array_equal(V.elem_dofs[v * spans + u], idx[v : v + degree + 1, u : u + degree + 1].ravel())


def test_expand_dofs_interleaves():
    np.testing.assert_array_equal(S.expand_dofs(np.array([[0, 3]]), 2), [[0, 1, 6, 7]])


def test_make_dofmap_sorted_and_lift():
    dm = S.make_dofmap(6, [4, 1], [2.0, 1.0])
    np.testing.assert_array_equal(dm.prescribed, [1, 4])
    np.testing.assert_array_equal(dm.free, [0, 2, 3, 5])
    np.testing.assert_allclose(dm.lift(np.array([9.0, 8, 7, 6])), [9, 1, 8, 7, 2, 6])


GEOMS = {"quad": lambda: G.rectangle(0, 0, 2, 1), "annulus": lambda: G.quarter_annulus(1.0, 4.0)}


@pytest.mark.parametrize("name", list(GEOMS))
def test_evaluate_partition_of_unity(name):
    V = S.function_space(GEOMS[name]().elevate(3).refine(2))
    bd = S.gauss_basis(V)
    np.testing.assert_allclose(bd.R.sum(-1), 1.0, atol=1e-13)
    np.testing.assert_allclose(bd.dR.sum(-1), 0.0, atol=1e-10)


def test_evaluate_exact_annulus_area_and_polynomial():
    V = S.function_space(G.quarter_annulus(1.0, 4.0).elevate(3).refine(3))
    np.testing.assert_allclose(S.gauss_basis(V).w.sum(), np.pi / 4 * 15, rtol=1e-12)
    V = S.function_space(G.rectangle(0, 0, 2, 3).elevate(3).refine(2))
    bd = S.gauss_basis(V)
    np.testing.assert_allclose((bd.w * bd.x[..., 0] ** 2 * bd.x[..., 1]).sum(), 12.0, rtol=1e-12)


@pytest.mark.parametrize("label,exact", [("outer", np.pi / 2 * 4), ("inner", np.pi / 2), ("sym_y", 3.0), ("sym_x", 3.0)])
def test_boundary_measure(label, exact):
    V = S.function_space(G.quarter_annulus(1.0, 4.0).elevate(3).refine(3))
    bd = S.boundary_basis(V, V.boundary(label))
    np.testing.assert_allclose(bd.w.sum(), exact, rtol=1e-12)


def test_normals_outward_and_divergence_theorem():
    V = S.function_space(G.quarter_annulus(1.0, 4.0).elevate(3).refine(2))
    for label, sign in (("outer", 1.0), ("inner", -1.0)):
        bd = S.boundary_basis(V, V.boundary(label))
        x = bd.x.reshape(-1, 2)
        n = bd.normal.reshape(-1, 2)
        np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-13)
        np.testing.assert_allclose((n * x / np.linalg.norm(x, axis=1, keepdims=True)).sum(1), sign, atol=1e-12)
    for label, exp in (("sym_y", [0.0, -1.0]), ("sym_x", [-1.0, 0.0])):
        n = S.boundary_basis(V, V.boundary(label)).normal.reshape(-1, 2)
        np.testing.assert_allclose(n, np.broadcast_to(exp, n.shape), atol=1e-13)
    area = S.gauss_basis(V).w.sum()
    flux = 0.0
    for label in ("outer", "inner", "sym_y", "sym_x"):
        bd = S.boundary_basis(V, V.boundary(label))
        flux += (bd.w * (bd.x * bd.normal).sum(-1) / 2).sum()
    np.testing.assert_allclose(flux, area, rtol=1e-11)


@pytest.mark.parametrize("name", list(GEOMS))
def test_affine_fields_reproduced(name):
    V = S.function_space(GEOMS[name]().elevate(3).refine(2))
    bd = S.gauss_basis(V)
    for comp in range(2):
        local = V.cpts[:, comp][bd.dofs]
        first = np.einsum("eqdn,en->eqd", bd.dR, local)
        exp = np.zeros_like(first)
        exp[..., comp] = 1.0
        np.testing.assert_allclose(first, exp, atol=1e-10)


def test_cube_volume_and_partition():
    V = S.function_space(G.cuboid([0, 0, 0], [1, 2, 3]).elevate(2).refine(1))
    bd = S.gauss_basis(V)
    np.testing.assert_allclose(bd.R.sum(-1), 1.0, atol=1e-13)
    np.testing.assert_allclose(bd.w.sum(), 6.0, rtol=1e-12)
