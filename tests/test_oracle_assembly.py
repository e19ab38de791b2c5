"""Pins the oracle's CSR pattern, scatter-add assembly and load vector against
brute force (reference tests/test_advanced.py:40-43,135-151 and
tests/test_galerkin.py:409-447)."""

import numpy as np

from oracle import assembly as A
from oracle import forms as F
from oracle import geometry as G
from oracle import space as S


def _system(vec=1):
    if vec == 1:
        V = S.function_space(G.rectangle(0, 0, 1, 1).elevate(3).refine(2))
        bcs = [F.DirichletBC(0.0, lambda x: np.ones(len(x), bool))]
    else:
        V = S.function_space(G.quarter_annulus(1.0, 2.0).elevate(2).refine(2), vec=2)
        bcs = [F.DirichletBC(0.0, "inner")]
    presc, vals = F.resolve_dirichlet(V, bcs)
    return V, S.make_dofmap(V.n_dofs, presc, vals)


def test_pattern_equals_brute_force_dense_pattern():
    for vec in (1, 2):
        V, dm = _system(vec)
        pat = A.csr_pattern(V.elem_dofs, V.n_dofs, V.vec, dm.free)
        n = len(dm.free)
        to_free = np.full(V.n_dofs, -1)
        to_free[dm.free] = np.arange(n)
        D = np.zeros((n, n), bool)
        for ed in S.expand_dofs(V.elem_dofs, V.vec):
            f = to_free[ed]
            f = f[f >= 0]
            D[np.ix_(f, f)] = True
        rows, cols = np.nonzero(D)
        np.testing.assert_array_equal(pat.row_indices, rows)
        np.testing.assert_array_equal(pat.column_indices, cols)
        np.testing.assert_array_equal(pat.indptr, np.concatenate([[0], np.cumsum(D.sum(1))]))
        keys = pat.row_indices.astype(np.int64) * n + pat.column_indices
        assert np.all(np.diff(keys) > 0) and pat.indptr[-1] == pat.nnz


def test_pattern_independent_of_chunk_size():
    V, dm = _system(2)
    a = A.csr_pattern(V.elem_dofs, V.n_dofs, V.vec, dm.free)
    b = A.csr_pattern(V.elem_dofs, V.n_dofs, V.vec, dm.free, chunk=3000)
    for x, y in ((a.indptr, b.indptr), (a.column_indices, b.column_indices), (a.scatter, b.scatter)):
        np.testing.assert_array_equal(x, y)


def test_assembly_equals_dense_add_at():
    V, dm = _system(2)
    bd = S.gauss_basis(V)
    Ke = F.elasticity_tangents(bd, {"E": 1.0, "nu": 0.3}, 2, "strain")
    pat = A.csr_pattern(V.elem_dofs, V.n_dofs, V.vec, dm.free)
    data = A.assemble_csr(pat, Ke)
    Kfull = np.zeros((V.n_dofs, V.n_dofs))
    gd = S.expand_dofs(V.elem_dofs, 2)
    for e in range(V.n_elems):
        np.add.at(Kfull, (gd[e][:, None], gd[e][None, :]), Ke[e])
    Kff = Kfull[np.ix_(dm.free, dm.free)]
    np.testing.assert_allclose(A.to_scipy(pat, data).toarray(), Kff, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(Kff, Kff.T, rtol=1e-13, atol=1e-14)


def test_load_integrates_the_source_and_lifting():
    V, dm = _system(1)
    bd = S.gauss_basis(V)
    gd = S.expand_dofs(V.elem_dofs, 1)
    Fe = F.source_load(bd, lambda x, p: 1.0 + x[0] * x[1], None, 1)
    load = A.global_load(V.n_dofs, gd, Fe)
    # partition of unity: the loads sum to the integral of f over the square
    np.testing.assert_allclose(load.sum(), 1.25, rtol=1e-12)
    Ke = F.poisson_tangents(bd, None)
    # lifting of a constant Dirichlet value is -K_fp g = K_ff 1 * g for a
    # constant field (rows of the full stiffness sum to zero)
    dm2 = S.make_dofmap(V.n_dofs, dm.prescribed, np.full(len(dm.prescribed), 2.0))
    lift = A.lifting(V.n_dofs, gd, Ke, dm2)[dm2.free]
    pat = A.csr_pattern(V.elem_dofs, V.n_dofs, 1, dm2.free)
    Kff = A.to_scipy(pat, A.assemble_csr(pat, Ke))
    np.testing.assert_allclose(lift, -(Kff @ np.full(len(dm2.free), 2.0)), rtol=1e-11, atol=1e-12)
