"""The element-list path (the padded-element contract and chunked assembly)
against the oracle and the tensor path:

* CSR pattern bit-exact with the oracle's sorted-unique-keys pattern, and
  the slot map equal to the oracle's scatter map after undoing the local
  permutation, for spaces padded with the sentinel dof (three extra
  functions per element, zero extraction rows) and locally permuted;
* CSR values and load within 1e-12 of the oracle's (FP64 atomics: the
  summation order differs), padded or not, whole or in chunks;
* solutions independent of the chunk size (the reference's chunk
  invariance, tests/test_galerkin.py:409-447) and equal to the tensor path's.
"""

import numpy as np
import pytest
import torch

import jaxiga_amd as jx
from jaxiga_amd import elements as EL
from jaxiga_amd import grf
from jaxiga_amd import methods as M
from oracle import assembly as OA
from oracle import forms as OF
from tests import cases as C
from tests.test_gpu_parity import _oracle_system, rel

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "square":
        return C.poisson_case("square", 3, 3)
    if name == "annulus":
        return C.poisson_case("annulus", 2, 2, a0=lambda x, p: 2.0 + x[1] * x[0])
    if name == "cube":
        case = C.poisson_case("cube", 2, 2, source=lambda x, p: 1.0 + 0 * x[0])
        ctx = M.build_context(case.problem)
        a = grf.gauss_point_field(ctx, 7, 0.2, 1.0).exp_()
        case.params = {"a0": jx.PointField(a)}
        case.oparams = {"a0": OF.PointField(a.cpu().numpy())}
        return case
    if name == "elasticity":
        c = C.elasticity_case(2, 2, "stress")
        c.problem = jx.LinearElasticity(c.space, plane="stress", dirichlet=[jx.DirichletBC(0.0, where="inner")],
                                        source=lambda x, p: C.xp(x).stack([0.1 + 0 * x[0], -1.0 + x[0]]))
        c.oneumann = []
        c.source = lambda x, p: np.stack([0.1 + 0 * x[0], -1.0 + x[0]])
        return c
    raise ValueError(name)


def _oracle(case):
    saved = case.params
    case.params = getattr(case, "oparams", case.params)
    try:
        return _oracle_system(case)
    finally:
        case.params = saved


@pytest.mark.parametrize("name", ["square", "annulus", "cube", "elasticity"])
@pytest.mark.parametrize("padded", [False, True])
def test_element_path_matches_oracle(gpu, name, padded):
    case = _case(name)
    print(f"[elements] {name} padded={padded}", flush=True)
    dm, pat, data, b = _oracle(case)
    V = case.space
    es = EL.ElementSpace.from_space(V, n_local_max=V.n_local + 3 if padded else None,
                                    permute_seed=11 if padded else None)
    prob = case.problem.with_space(es)
    ctx = EL.build_context(prob, es, case.params)
    np.testing.assert_array_equal(ctx.dofmap.free, dm.free)
    gp = ctx.pattern
    np.testing.assert_array_equal(gp.indptr, pat.indptr)
    np.testing.assert_array_equal(gp.column_indices, pat.column_indices)
    np.testing.assert_array_equal(gp.row_indices, pat.row_indices)
    # the slot map, in the tensor space's local order, equals the oracle's
    sc = gp.scatter_d.cpu().numpy()
    vec = V.vec
    ed = V.elem_dofs
    for e in range(0, V.n_elems, max(1, V.n_elems // 7)):
        loc = {int(d): k for k, d in enumerate(es.elem_dofs[e]) if d < es.n_scalar_basis}
        order = np.array([loc[int(d)] for d in ed[e]])
        rows = (order[:, None] * vec + np.arange(vec)).ravel()
        np.testing.assert_array_equal(sc[e][np.ix_(rows, rows)],
                                      pat.scatter[e].reshape(len(rows), len(rows)))
        pad = [k for k, d in enumerate(es.elem_dofs[e]) if d == es.n_scalar_basis]
        for k in pad:
            assert (sc[e][k * vec:(k + 1) * vec] == gp.nnz).all()
    for chunk in (None, 5):
        vals, rhs = EL.assemble(ctx, es, case.params, chunk=chunk)
        assert rel(vals.cpu(), data) < 1e-12
        assert rel(rhs.cpu(), b) < 1e-12


def test_chunked_solve_matches_tensor_solve(gpu):
    case = C.poisson_case("square", 3, 4, a0=lambda x, p: 1.0 + x[0] * x[1])
    lin = jx.LinearOptions(method="cg", tol=1e-13)
    ref = jx.solve(case.problem, params=case.params, linear=lin)
    for chunk in (1, 7, 100, 10**6):
        sol = jx.solve(case.problem, params=case.params, linear=lin, chunk=chunk)
        assert sol.stats["converged"] and sol.stats["chunk"] == chunk
        assert float((sol.u - ref.u).abs().max()) <= 1e-11 * float(ref.u.abs().max())
    with pytest.raises(ValueError, match="chunk"):
        jx.solve(case.problem, params=case.params, linear=lin, chunk=0)


def test_chunked_elasticity_with_neumann_matches_oracle(gpu):
    case = C.elasticity_case(2, 2, "strain")
    dm, pat, data, b = _oracle(case)
    sol = jx.solve(case.problem, params=case.params, linear=jx.LinearOptions(method="cg", tol=1e-13), chunk=9)
    K = OA.to_scipy(pat, data)
    import scipy.sparse.linalg as spla

    x = spla.spsolve(K.tocsc(), b)
    assert rel(sol.u.cpu().numpy()[dm.free], x) < 1e-9


def test_padded_space_solve_matches_tensor_solve(gpu):
    case = C.poisson_case("annulus", 3, 2)
    lin = jx.LinearOptions(method="cg", tol=1e-13)
    ref = jx.solve(case.problem, params=case.params, linear=lin)
    es = EL.ElementSpace.from_space(case.space, n_local_max=case.space.n_local + 5, permute_seed=3)
    sol = jx.solve(case.problem.with_space(es), params=case.params, linear=lin)
    assert sol.stats["converged"]
    assert float((sol.u - ref.u).abs().max()) <= 1e-11 * float(ref.u.abs().max())
