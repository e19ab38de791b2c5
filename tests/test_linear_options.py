"""LinearOptions validation (CPU): the reference's checks
(tests/test_solver_robustness.py:13-22, tests/test_advanced.py:234-238,286-288)
and the multigrid preconditioner this build adds."""

import numpy as np
import pytest

from jaxiga_amd.linear import LinearOptions


@pytest.mark.parametrize("tol", [0, -1e-6, np.inf, np.nan])
def test_linear_options_reject_invalid_tolerance(tol):
    with pytest.raises(ValueError, match="tol"):
        LinearOptions(tol=tol)


@pytest.mark.parametrize("maxiter", [0, -1, 1.5, True])
def test_linear_options_reject_invalid_iteration_limit(maxiter):
    with pytest.raises(ValueError, match="maxiter"):
        LinearOptions(maxiter=maxiter)


def test_superlu_ordering_and_matrix_free_method_are_validated():
    with pytest.raises(ValueError, match="SuperLU ordering"):
        LinearOptions(superlu_ordering="not-an-ordering")
    with pytest.raises(ValueError, match="iterative"):
        LinearOptions(method="scipy", matrix_free=True)
    with pytest.raises(ValueError, match="unknown linear method"):
        LinearOptions(method="lu")


def test_multigrid_preconditioner_needs_assembled_cg():
    LinearOptions(method="cg", preconditioner="multigrid", tol=1e-10)
    for kwargs in ({"method": "dense"}, {"method": "gmres"}, {"method": "cg", "matrix_free": True}):
        with pytest.raises(ValueError, match="multigrid"):
            LinearOptions(preconditioner="multigrid", **kwargs)
    with pytest.raises(ValueError, match="unknown preconditioner"):
        LinearOptions(method="cg", preconditioner="ilu")


def _unit_poisson():
    import jaxiga_amd as jx

    V = jx.FunctionSpace(jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(2).refine(2))
    return V, jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))])


def test_default_linear_follows_the_cg_crossover():
    """reference methods/galerkin.py:248-272: CG (Jacobi, tol 1e-12) on an
    accelerator for declared-SPD linear problems with at least CG_CROSSOVER
    free dofs, the default LinearOptions() (SuperLU) below it."""
    import jaxiga_amd as jx
    from jaxiga_amd.methods import CG_CROSSOVER, default_linear

    assert CG_CROSSOVER == 3000
    V, prob = _unit_poisson()
    assert default_linear(prob, CG_CROSSOVER - 1) == LinearOptions()
    for n in (CG_CROSSOVER, None):
        big = default_linear(prob, n)
        assert (big.method, big.preconditioner, big.tol) == ("cg", "jacobi", 1e-12)

    class Custom(jx.Poisson):  # overrides the energy without re-declaring SPD
        def energy(self, grad_u, u, x, params):
            return 0.0

    class Declared(Custom):
        is_positive_definite = True

    bc = [jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))]
    assert default_linear(Custom(V, dirichlet=bc), 10**6) == LinearOptions()
    assert default_linear(Declared(V, dirichlet=bc), 10**6).method == "cg"


def test_scipy_method_is_host_superlu():
    """method='scipy' factors on the host with SuperLU like the reference
    (solvers/linear.py:221-250), at any size; method='dense' keeps the
    device LU's size limit."""
    import scipy.sparse
    import torch

    from jaxiga_amd import linear as LIN

    n = 30
    rng = np.random.default_rng(0)
    A = scipy.sparse.random(n, n, density=0.2, random_state=1)
    A = (A + A.T + n * scipy.sparse.eye(n)).tocsr()
    A.sort_indices()
    coo = A.tocoo()
    order = np.lexsort((coo.col, coo.row))
    pat = LIN.CSRPattern(torch.as_tensor(A.indptr.astype(np.int32)),
                         torch.as_tensor(coo.col[order].astype(np.int32)),
                         torch.as_tensor(coo.row[order].astype(np.int32)), n)
    data = torch.as_tensor(A.data)
    b = torch.as_tensor(rng.normal(size=n))
    for ordering in ("auto", "COLAMD", "NATURAL"):
        u, info = LIN._solve_impl(LinearOptions(superlu_ordering=ordering), pat, data, b)
        assert info["iterations"] is None
        np.testing.assert_allclose(A @ u.numpy(), b.numpy(), rtol=0, atol=1e-12)
    big = LIN.CSRPattern(torch.zeros(LIN.DIRECT_LIMIT + 2, dtype=torch.int32),
                         torch.zeros(0, dtype=torch.int32), torch.zeros(0, dtype=torch.int32),
                         LIN.DIRECT_LIMIT + 1)
    with pytest.raises(ValueError, match="dense solves are limited"):
        LIN._solve_impl(LinearOptions(method="dense"), big, torch.zeros(0, dtype=torch.float64),
                        torch.zeros(LIN.DIRECT_LIMIT + 1, dtype=torch.float64))


def test_point_callables_written_per_point_are_vectorised():
    """A callable written for one point (a reduction over x) is evaluated per
    point, as the reference's per-point calls under vmap do; constants stay
    constants."""
    import torch

    from jaxiga_amd.methods import call_points

    x = torch.as_tensor(np.random.default_rng(3).uniform(size=(6, 4, 2)))
    out = call_points(lambda y, p: 1 + y.sum(), x)
    np.testing.assert_allclose(out[..., 0].numpy(), 1 + x.sum(-1).numpy(), rtol=1e-15)
    out = call_points(lambda y, p: torch.linalg.norm(y), x)
    np.testing.assert_allclose(out[..., 0].numpy(), x.norm(dim=-1).numpy(), rtol=1e-15)
    out = call_points(lambda y, p: 1.0 + y[0] * y[1], x)
    np.testing.assert_allclose(out[..., 0].numpy(), (1.0 + x[..., 0] * x[..., 1]).numpy(), rtol=1e-15)
    assert float(call_points(lambda y, p: 2.5, x).unique()) == 2.5
    vec = call_points(lambda y, p: torch.tensor([0.0, -1.0], dtype=torch.float64), x, vec=2)
    assert vec.shape == (6, 4, 2) and float(vec[..., 1].unique()) == -1.0


def test_element_space_validates_the_padding_contract():
    """reference space/function_space.py:23-38: padding entries are the
    sentinel n_scalar_basis and their extraction rows are zero."""
    import jaxiga_amd as jx
    from jaxiga_amd.elements import ElementSpace

    V, _ = _unit_poisson()
    es = ElementSpace.from_space(V, n_local_max=V.n_local + 2, permute_seed=0)
    assert es.n_local_max == V.n_local + 2
    assert (es.elem_dofs == V.n_scalar_basis).sum() == 2 * V.n_elems
    bad = es.extraction.copy()
    bad[es.elem_dofs == V.n_scalar_basis] = 1.0
    with pytest.raises(ValueError, match="zero"):
        ElementSpace(degree=V.degree, cpts=V.cpts, wgts=V.wgts, elem_dofs=es.elem_dofs, extraction=bad,
                     elem_vertex=V.elem_vertex)
    with pytest.raises(ValueError, match="n_local_max"):
        ElementSpace.from_space(V, n_local_max=V.n_local - 1)
    del jx
