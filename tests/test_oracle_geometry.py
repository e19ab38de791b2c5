"""Pins the oracle's knot insertion and degree elevation (geometry invariance,
reference tests/test_patch.py:29-38), the closed-form Bezier elevation, and
the product's host refinement against the oracle."""

import numpy as np
import pytest
from math import comb

import jaxiga_amd as jx
from oracle import geometry as G

RNG = np.random.default_rng(7)
PARAMS2 = RNG.uniform(0, 1, size=(40, 2))
PARAMS3 = RNG.uniform(0, 1, size=(20, 3))

PATCHES = {
    "square": (lambda: G.rectangle(0, 0, 2, 1), PARAMS2),
    "annulus": (lambda: G.quarter_annulus(1.0, 4.0), PARAMS2),
    "cube": (lambda: G.cuboid([0, 0, 0], [1, 2, 3]), PARAMS3),
}


@pytest.mark.parametrize("name", list(PATCHES))
def test_elevate_and_refine_leave_the_geometry_unchanged(name):
    make, params = PATCHES[name]
    base = make()
    x0 = base.evaluate(params)
    for patch in (base.elevate(3), base.elevate(3).refine(2), base.refine(1).elevate(4)):
        np.testing.assert_allclose(patch.evaluate(params), x0, rtol=1e-11, atol=1e-12)


def test_single_segment_elevation_is_the_bernstein_formula():
    p, t = 2, 2
    c = RNG.normal(size=(p + 1, 2))
    kv = np.array([0, 0, 0, 1, 1, 1.0])
    out, knots = G.elevate_1d(p, c, kv, t)
    q = p + t
    exp = np.zeros((q + 1, 2))
    for i in range(q + 1):
        for j in range(max(0, i - t), min(p, i) + 1):
            exp[i] += comb(p, j) * comb(t, i - j) / comb(q, i) * c[j]
    np.testing.assert_allclose(out, exp, rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(knots, [0] * (q + 1) + [1] * (q + 1))


def test_boehm_insertion_preserves_curve_and_matches_scipy():
    from scipy.interpolate import BSpline

    p = 3
    kv = np.array([0, 0, 0, 0, 0.4, 1, 1, 1, 1.0])
    c = RNG.normal(size=(5, 1))
    new_c, new_k = G.insert_knots_1d(p, c, kv, [0.2, 0.4, 0.7])
    x = np.linspace(0, 1, 33)
    np.testing.assert_allclose(BSpline(new_k, new_c[:, 0], p)(x), BSpline(kv, c[:, 0], p)(x), atol=1e-13)


def test_refine_counts():
    P = G.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(3).refine(5)
    assert P.n_cp_per_dir == (35, 35)


@pytest.mark.parametrize("geom", ["square", "annulus", "cube"])
def test_product_host_refinement_matches_oracle(geom):
    if geom == "square":
        P, O = jx.primitives.rectangle(0, 0, 2, 1), G.rectangle(0, 0, 2, 1)
    elif geom == "annulus":
        P, O = jx.primitives.quarter_annulus(1.0, 4.0), G.quarter_annulus(1.0, 4.0)
    else:
        P, O = jx.primitives.cuboid([0, 0, 0], [1, 2, 3]), G.cuboid([0, 0, 0], [1, 2, 3])
    P, O = P.elevate(3).refine(2), O.elevate(3).refine(2)
    assert P.n_cp_per_dir == O.n_cp_per_dir
    for a, b in zip(P.knots, O.knots):
        np.testing.assert_array_equal(np.asarray(a), b)
    np.testing.assert_allclose(P.ctrl_pts, O.ctrl_pts, rtol=0, atol=1e-13)
    np.testing.assert_allclose(P.weights, O.weights, rtol=0, atol=1e-13)
