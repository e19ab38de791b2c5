"""Pins the multigrid oracle: the Boehm transfer reproduces the coarse
B-splines, the Galerkin operator equals the coarse-space stiffness assembled
directly, the stencil smoother and its gradient against a dense statement,
adjointness and finite differences, the V-cycle is symmetric and makes FCG
converge to the sparse direct solution. Also checks the product's host-side
coarsening and collocation transfer against the oracle."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import assembly as A
from oracle import forms as F
from oracle import geometry as G
from oracle import multigrid as M
from oracle import space as S


def _poisson(refine, p=3, a0=1.0):
    patch = G.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(p).refine(refine)
    V = S.function_space(patch)
    basis = S.gauss_basis(V)
    params = {"a0": a0} if not callable(a0) else {"a0": F.PointField(a0(basis.x))}
    Ke = F.poisson_tangents(basis, params)
    presc, _ = F.resolve_dirichlet(V, [F.DirichletBC(0.0, lambda x: np.ones(len(x), bool))])
    dm = S.make_dofmap(V.n_dofs, presc, np.zeros(len(presc)))
    pat = A.csr_pattern(V.elem_dofs, V.n_dofs, 1, dm.free)
    K = A.to_scipy(pat, A.assemble_csr(pat, Ke)).tocsr()
    return K, patch, dm


def test_refinement_matrix_reproduces_coarse_splines():
    p = 3
    fine = np.concatenate([np.zeros(p + 1), np.arange(1, 8) / 8, np.ones(p + 1)])
    coarse = M.coarsen_knots(fine, p)
    np.testing.assert_allclose(coarse, np.concatenate([np.zeros(p + 1), np.arange(1, 4) / 4, np.ones(p + 1)]))
    T = M.refinement_matrix(coarse, fine, p)
    x = np.random.default_rng(0).random(200)
    np.testing.assert_allclose(G.basis_matrix(fine, p, x) @ T, G.basis_matrix(coarse, p, x), atol=1e-13)
    assert M.coarsen_knots(np.concatenate([np.zeros(4), [1 / 3, 2 / 3], np.ones(4)]), 3) is None


def test_product_transfer_matches_oracle():
    from jaxiga_amd.geometry import change_of_basis
    from jaxiga_amd.multigrid import coarsen_knots

    for p, n in [(2, 8), (3, 16), (3, 6)]:
        u = np.linspace(0, 1, n + 1) ** 1.3  # non-uniform
        fine = np.concatenate([np.zeros(p), u, np.ones(p)])
        c = coarsen_knots(fine, p)
        np.testing.assert_array_equal(c, M.coarsen_knots(fine, p))
        np.testing.assert_allclose(change_of_basis(c, p, fine, p), M.refinement_matrix(c, fine, p), atol=1e-12)


def test_galerkin_equals_direct_coarse_stiffness():
    K8, patch8, dm8 = _poisson(3)
    K16, patch16, dm16 = _poisson(4)
    levels = M.hierarchy(K16, patch16.knots, 3, dm16.prescribed, max_coarse=1)
    assert [lv["K"].shape[0] for lv in levels[:2]] == [K16.shape[0], K8.shape[0]]
    assert (levels[0]["h"], levels[0]["w"]) == (17, 17)
    D = (levels[1]["K"] - K8).toarray()
    assert np.abs(D).max() <= 1e-12 * np.abs(K8.toarray()).max()


def _random_stencil(rng, ks, n):
    W = rng.normal(size=(ks * ks, n)) * 0.1
    W[(ks * ks) // 2] += 1.0
    return W


def test_stencil_apply_dense_statement_and_adjoint():
    rng = np.random.default_rng(3)
    h, w, ks = 6, 5, 5
    n = h * w
    W = _random_stencil(rng, ks, n)
    dsq = rng.random(n) + 0.5
    T = np.zeros((n, n))
    r = ks // 2
    for o, (dy, dx) in enumerate([(a, b) for a in range(-r, r + 1) for b in range(-r, r + 1)]):
        for i in range(n):
            y, x = divmod(i, w)
            if 0 <= y + dy < h and 0 <= x + dx < w:
                T[i, (y + dy) * w + x + dx] = W[o, i]
    S_dense = np.diag(dsq) @ T @ np.diag(dsq)
    v, u = rng.normal(size=n), rng.normal(size=n)
    np.testing.assert_allclose(M.stencil_apply(W, dsq, v, h, w, ks), S_dense @ v, atol=1e-13)
    np.testing.assert_allclose(M.stencil_apply(W, dsq, u, h, w, ks, transpose=True), S_dense.T @ u, atol=1e-13)


def test_energy_loss_gradient_finite_differences():
    rng = np.random.default_rng(4)
    K, patch, dm = _poisson(2)
    h = w = 5
    n = h * w
    assert K.shape[0] == n
    ks = 5
    W = _random_stencil(rng, ks, n)
    dsq = 1 / np.sqrt(K.diagonal())
    E = rng.normal(size=(3, n))
    R = (K @ E.T).T
    loss, dW = M.energy_loss(K, W, dsq, E, R, h, w, ks)
    for _ in range(6):
        o, i = rng.integers(0, ks * ks), rng.integers(0, n)
        Wp, Wm = W.copy(), W.copy()
        Wp[o, i] += 1e-6
        Wm[o, i] -= 1e-6
        fd = (M.energy_loss(K, Wp, dsq, E, R, h, w, ks)[0] - M.energy_loss(K, Wm, dsq, E, R, h, w, ks)[0]) / 2e-6
        assert abs(fd - dW[o, i]) <= 1e-6 * max(1.0, abs(fd))


def test_stencil_features_dense_statement():
    K, patch, dm = _poisson(2)
    h = w = 5
    Fe = M.stencil_features(K, h, w, 3)
    Kd = K.toarray()
    for i in range(h * w):
        for j in range(h * w):
            dr, dc = j // w - i // w, j % w - i % w
            if abs(dr) <= 3 and abs(dc) <= 3:
                assert Fe[(dr + 3) * 7 + dc + 3, i] == pytest.approx(Kd[i, j] / Kd[i, i], abs=1e-15)


def test_vcycle_symmetric_and_fcg_converges():
    rng = np.random.default_rng(5)
    coef = lambda x: np.exp(np.sin(6 * x[..., 0]) * np.cos(4 * x[..., 1]))
    K, patch, dm = _poisson(4, a0=coef)
    levels = M.hierarchy(K, patch.knots, 3, dm.prescribed)
    assert len(levels) == 3 and levels[-1]["K"].shape[0] <= 64
    stencils = [_random_stencil(rng, 7, lv["K"].shape[0]) * 0.5 for lv in levels[:-1]]
    a, c = rng.normal(size=K.shape[0]), rng.normal(size=K.shape[0])
    for st in (None, stencils):
        Va = M.vcycle(levels, st, a)
        Vc = M.vcycle(levels, st, c)
        assert abs(Va @ c - a @ Vc) <= 1e-10 * abs(Va @ c)
    b = rng.normal(size=K.shape[0])
    x, it = M.fcg(K, b, lambda r: M.vcycle(levels, None, r), tol=1e-10)
    xd = spla.spsolve(K.tocsc(), b)
    assert it < 40
    assert np.abs(x - xd).max() <= 1e-7 * np.abs(xd).max()
    # exact two-level solve: with a single level the V-cycle is the dense inverse
    one = M.hierarchy(K, patch.knots, 3, dm.prescribed, max_coarse=10**6)
    assert len(one) == 1
    np.testing.assert_allclose(M.vcycle(one, None, b), xd, rtol=1e-9, atol=1e-12)
