"""GPU parity of the matrix-free sum-factorised operator against the oracle's
assembled FP64 operator (reference tests/test_advanced.py:123-132,198-215,
240-269: K v to 1e-12, diagonal, solution, equal PCG iteration counts)."""

import numpy as np
import pytest
import torch

import jaxiga_amd as jx
from jaxiga_amd import _lib
from jaxiga_amd import matfree as MF
from jaxiga_amd import methods as M
from oracle import assembly as OA
from oracle import solver as OSV
from tests import cases as C
from tests.test_gpu_parity import _oracle_system

pytestmark = pytest.mark.gpu

MF_CASES = [
    lambda: C.poisson_case("square", 3, 3),
    lambda: C.poisson_case("square", 2, 3, a0=lambda x, p: 1.0 + x[0] * x[1]),
    lambda: C.poisson_case("annulus", 3, 2),
    lambda: C.poisson_case("annulus", 2, 2, a0=lambda x, p: 2.0 + x[1]),
    lambda: C.poisson_case("cube", 2, 2, source=lambda x, p: 1.0 + 0 * x[0]),
    lambda: C.poisson_case("cube", 1, 2, source=lambda x, p: 1.0 + 0 * x[0]),
]


def _params_for_oracle(case):
    a0 = case.params.get("a0")
    if callable(a0):
        return {"a0": lambda x, p: a0(x, p)}
    return case.params


@pytest.mark.parametrize("make", MF_CASES)
def test_matrix_free_apply_equals_assembled_product(gpu, make):
    case = make()
    print(f"[mf] {case.name}", flush=True)
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    v = np.random.default_rng(0).normal(size=K.shape[0])
    ctx = M.build_context(case.problem, case.params)
    op = MF.MatrixFreeOperator(ctx, case.params)
    y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
    ref = K @ v
    assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-12
    # single precision apply: FP32 rounding of the geometric factors and sums
    op32 = MF.MatrixFreeOperator(ctx, case.params, precision=_lib.IGA_FP32)
    y32 = op32.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
    assert np.abs(y32 - ref).max() / np.abs(ref).max() < 5e-6


@pytest.mark.parametrize("spread", [(1.0, 4e-16), (2.5, 0.0)])
def test_constant_weights_take_the_polynomial_apply(gpu, spread):
    """Weights equal up to rounding (knot insertion leaves weights of 1 a few
    ulps off) or all scaled by one constant give R_i = N_i, so the operator
    takes the polynomial apply; it must still equal the oracle's rational
    assembly on the same weights."""
    import dataclasses

    from oracle import space as OS

    scale, ulps = spread
    P, O = C.patches("square", 2, 3)
    w = O.weights * scale * (1.0 + ulps * np.where(np.arange(O.weights.size) % 2 == 0, 1.0, -1.0))
    assert np.ptp(w) > 0 or ulps == 0
    P = dataclasses.replace(P, weights=w)
    O = dataclasses.replace(O, weights=w.copy())
    a0 = lambda x, p: 1.0 + x[0] * x[1]  # noqa: E731
    V = jx.FunctionSpace(P)
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=C.everywhere)], source=C.sin_source)
    case = C.Case("poisson-square-p2-r3-w", "poisson", V, OS.function_space(O), prob,
                  [C.OF.DirichletBC(0.0, C.everywhere)], [], C.sin_source, {"a0": a0})
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    v = np.random.default_rng(1).normal(size=K.shape[0])
    ctx = M.build_context(case.problem, case.params)
    op = MF.MatrixFreeOperator(ctx, case.params)
    assert op.rational == 0
    y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
    ref = K @ v
    assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-12


def test_matrix_free_diagonal_equals_assembled_diagonal(gpu):
    case = C.poisson_case("cube", 2, 2, source=lambda x, p: 1.0 + 0 * x[0])
    dm, pat, data, b = _oracle_system(case)
    ctx = M.build_context(case.problem, case.params)
    d = M.tangent_diagonal(ctx, case.params).cpu().numpy()
    np.testing.assert_allclose(d, OA.to_scipy(pat, data).diagonal(), rtol=1e-12)


def test_matrix_free_solution_equals_dense(gpu):
    case = C.poisson_case("square", 2, 3)
    ref = jx.solve(case.problem, params=case.params, linear=jx.LinearOptions(method="dense"))
    mf = jx.solve(case.problem, params=case.params,
                  linear=jx.LinearOptions(method="cg", matrix_free=True, tol=1e-13))
    assert mf.stats["matrix_free"] and mf.stats["nnz"] is None and mf.stats["converged"]
    assert float((mf.u - ref.u).abs().max()) < 1e-12


@pytest.mark.parametrize("make", [MF_CASES[0], MF_CASES[2], MF_CASES[4]])
def test_matrix_free_takes_the_same_iteration_count(gpu, make):
    case = make()
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    _, it_ref = OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-10)
    ctx = M.build_context(case.problem, case.params)
    vals, rhs, diag = M.assemble(ctx, case.params, diagonal=True)
    assembled = jx.linear.cg_csr(ctx.assembly, vals, rhs, tol=1e-10)
    op = MF.MatrixFreeOperator(ctx, case.params)
    res = op.cg(rhs, tol=1e-10, dinv=MF.inverse_guard(diag))
    assert res.iterations == assembled.iterations
    assert abs(res.iterations - it_ref) <= 1
    np.testing.assert_allclose(res.x.cpu().numpy(), assembled.x.cpu().numpy(), rtol=0, atol=1e-9 * float(assembled.x.abs().max()))
