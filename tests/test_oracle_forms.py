"""Pins the oracle's element kernels and boundary-condition resolution.

The kernels must equal the Hessian of the element energy (the check of the
reference's tests/test_advanced.py:154-195): the energy is evaluated here from
its pointwise density (1/2 a |grad u|^2, 1/2 eps:C:eps with engineering
strains) and differentiated by exact second differences -- the energy is
quadratic, so unit steps are exact up to rounding.
"""

import numpy as np
import pytest

from oracle import forms as F
from oracle import geometry as G
from oracle import space as S


def _hessian(energy, n):
    H = np.zeros((n, n))
    E = np.eye(n)
    for i in range(n):
        for j in range(n):
            H[i, j] = (energy(E[i] + E[j]) - energy(E[i] - E[j]) - energy(-E[i] + E[j])
                       + energy(-E[i] - E[j])) / 4.0
    return H


def test_poisson_kernel_is_the_energy_hessian():
    V = S.function_space(G.rectangle(0, 0, 1, 1).elevate(2).refine(1))
    bd = S.gauss_basis(V)
    a0 = lambda x, p: 1.0 + x[0]  # noqa: E731
    K = F.poisson_tangents(bd, {"a0": a0})
    a = 1.0 + bd.x[..., 0]
    for e in (0, 3):
        def energy(u):
            g = np.einsum("qdn,n->qd", bd.dR[e], u)
            return np.sum(bd.w[e] * 0.5 * a[e] * (g**2).sum(-1))
        np.testing.assert_allclose(K[e], _hessian(energy, V.n_local), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("plane", ["stress", "strain"])
def test_elasticity_kernel_is_the_energy_hessian(plane):
    V = S.function_space(G.quarter_annulus(1.0, 2.0).elevate(2).refine(1), vec=2)
    bd = S.gauss_basis(V)
    params = {"E": 2.0, "nu": 0.27}
    K = F.elasticity_tangents(bd, params, 2, plane)
    C = F.constitutive(2.0, 0.27, 2, plane)
    for e in (0, 2):
        def energy(u):
            uu = u.reshape(V.n_local, 2)
            g = np.einsum("qdn,nc->qcd", bd.dR[e], uu)
            eps = np.stack([g[:, 0, 0], g[:, 1, 1], g[:, 0, 1] + g[:, 1, 0]], 1)
            return np.sum(bd.w[e] * 0.5 * np.einsum("qi,ij,qj->q", eps, C, eps))
        np.testing.assert_allclose(K[e], _hessian(energy, 2 * V.n_local), rtol=1e-11, atol=1e-12)


def test_elasticity_3d_kernel_is_the_energy_hessian():
    V = S.function_space(G.cuboid([0, 0, 0], [1, 2, 1]).elevate(1).refine(1), vec=3)
    bd = S.gauss_basis(V)
    K = F.elasticity_tangents(bd, {"E": 1.0, "nu": 0.3}, 3)
    C = F.constitutive(1.0, 0.3, 3)
    e = 1

    def energy(u):
        g = np.einsum("qdn,nc->qcd", bd.dR[e], u.reshape(V.n_local, 3))
        eps = np.stack([g[:, 0, 0], g[:, 1, 1], g[:, 2, 2], g[:, 1, 2] + g[:, 2, 1],
                        g[:, 0, 2] + g[:, 2, 0], g[:, 0, 1] + g[:, 1, 0]], 1)
        return np.sum(bd.w[e] * 0.5 * np.einsum("qi,ij,qj->q", eps, C, eps))

    np.testing.assert_allclose(K[e], _hessian(energy, 3 * V.n_local), rtol=1e-11, atol=1e-12)


def test_plane_strain_matrix_closed_form():
    E, nu = 1.0, 0.3
    C = F.plane_strain(E, nu)
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    np.testing.assert_allclose(C, [[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]], rtol=1e-14)


def test_dirichlet_resolution_sorted_components_and_conflicts():
    V = S.function_space(G.quarter_annulus(1.0, 2.0).elevate(2).refine(1), vec=2)
    dofs, vals = F.resolve_dirichlet(V, [F.DirichletBC(0.0, "sym_y", component=1),
                                         F.DirichletBC(0.0, "sym_x", component=0)])
    assert np.all(np.diff(dofs) > 0)
    assert set(dofs % 2) == {0, 1}
    exp = np.sort(np.concatenate([2 * V.boundary("sym_y").dofs + 1, 2 * V.boundary("sym_x").dofs]))
    np.testing.assert_array_equal(dofs, exp)
    with pytest.raises(ValueError, match="conflicting"):
        F.resolve_dirichlet(V, [F.DirichletBC(0.0, "inner"), F.DirichletBC(1.0, "sym_y")])


def test_neumann_predicate_selects_elements_by_side_centre():
    V = S.function_space(G.rectangle(0, 0, 1, 1).elevate(2).refine(2))
    bset = F.resolve_neumann(V, lambda c: c[:, 0] > 1 - 1e-9)
    np.testing.assert_array_equal(np.sort(bset.elems), np.sort(V.boundary("right").elems))
