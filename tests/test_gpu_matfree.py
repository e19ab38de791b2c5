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


@pytest.mark.parametrize("ref", [3, 4])
def test_matrix_free_3d_grf_coefficient(gpu, ref):
    """The unit-cube workload's input path: a = exp(GRF l=0.1, sigma=1) at the
    27 Gauss points of every p=2 element, given as a PointField, folded into
    the 3-D geometric factors; FP64 and FP32 apply against the oracle's
    assembled FP64 operator on the same coefficient values, and the
    matrix-free solve against the oracle's direct solution."""
    import scipy.sparse.linalg as spla

    from jaxiga_amd import grf
    from oracle import forms as OF
    from oracle import grf as OGRF
    from oracle import space as OS

    case = C.poisson_case("cube", 2, ref, source=lambda x, p: 1.0 + 0 * x[0])
    print(f"[mf-grf] {case.name}", flush=True)
    ctx = M.build_context(case.problem)
    a = grf.gauss_point_field(ctx, 20260901, 0.1, 1.0).exp_()
    a_host = a.cpu().numpy()
    # the coefficient against the oracle's own GRF at the oracle's Gauss points
    xo = OS.gauss_basis(case.ospace).x.reshape(-1, 3)
    om, ph = OGRF.rff_params(20260901, 3, 0.1)
    pick = np.random.default_rng(ref).choice(len(xo), 3000, replace=False)
    a_ref = np.exp(OGRF.rff_field(xo[pick], om, ph, 1.0))
    assert np.abs(a_host.reshape(-1)[pick] - a_ref).max() <= 1e-12 * np.abs(a_ref).max()
    case.params = {"a0": OF.PointField(a_host)}
    params = {"a0": jx.PointField(a)}
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    v = np.random.default_rng(2).normal(size=K.shape[0])
    ref_y = K @ v
    for prec, tol in ((_lib.IGA_FP64, 1e-12), (_lib.IGA_FP32, 5e-6)):
        op = MF.MatrixFreeOperator(ctx, params, precision=prec)
        y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
        err = np.abs(y - ref_y).max() / np.abs(ref_y).max()
        print(f"[mf-grf] precision {prec}: K v rel err {err:.2e}", flush=True)
        assert err < tol
    sol = jx.solve(case.problem, params=params,
                   linear=jx.LinearOptions(method="cg", matrix_free=True, tol=1e-12))
    assert sol.stats["converged"]
    x_ref = spla.spsolve(K.tocsc(), b)
    u_free = sol.u.cpu().numpy()[dm.free]
    assert np.abs(u_free - x_ref).max() <= 1e-9 * np.abs(x_ref).max()


def test_matrix_free_3d_rational_weights_and_coefficient(gpu):
    """The 3-D rational apply (quotient-rule terms) with a non-symmetric
    coefficient: cuboid p=2 whose weights are perturbed by up to +-30 %, a0 =
    1 + x0 + 2 x1 x2; K v against the oracle's assembled operator."""
    import dataclasses

    from oracle import space as OS

    P, O = C.patches("cube", 2, 2)
    rng = np.random.default_rng(4)
    w = O.weights * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, size=O.weights.shape))
    P = dataclasses.replace(P, weights=w)
    O = dataclasses.replace(O, weights=w.copy())
    a0 = lambda x, p: 1.0 + x[0] + 2.0 * x[1] * x[2]  # noqa: E731
    src = lambda x, p: 1.0 + 0 * x[0]  # noqa: E731
    V = jx.FunctionSpace(P)
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=C.everywhere)], source=src)
    case = C.Case("poisson-cube-p2-r2-rational", "poisson", V, OS.function_space(O), prob,
                  [C.OF.DirichletBC(0.0, C.everywhere)], [], src, {"a0": a0})
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    v = np.random.default_rng(5).normal(size=K.shape[0])
    ref_y = K @ v
    ctx = M.build_context(case.problem, case.params)
    for prec, tol in ((_lib.IGA_FP64, 1e-12), (_lib.IGA_FP32, 5e-6)):
        op = MF.MatrixFreeOperator(ctx, case.params, precision=prec)
        assert op.rational == 1
        y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
        err = np.abs(y - ref_y).max() / np.abs(ref_y).max()
        print(f"[mf-rational] precision {prec}: K v rel err {err:.2e}", flush=True)
        assert err < tol


def _bf16_round(a):
    """FP64 -> FP32 (nearest) -> BF16 (nearest even), as FP64 values."""
    u = np.asarray(a, np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _grf_cube(ref, seed=20260901):
    from jaxiga_amd import grf
    from oracle import forms as OF

    case = C.poisson_case("cube", 2, ref, source=lambda x, p: 1.0 + 0 * x[0])
    ctx = M.build_context(case.problem)
    a = grf.gauss_point_field(ctx, seed, 0.1, 1.0).exp_()
    case.params = {"a0": OF.PointField(a.cpu().numpy())}
    return case, ctx, a


def test_bf16_rounding_helper_is_round_to_nearest_even():
    """The test's own BF16 statement against values whose rounding is known."""
    ulp = 2.0 ** -7  # BF16 spacing in [1, 2)
    vals = np.array([1.0, 1.0 + ulp / 2, 1.0 + 1.5 * ulp, 1.0 + ulp / 2 + 2 ** -20, -3.0])
    np.testing.assert_array_equal(_bf16_round(vals), [1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -3.0])


@pytest.mark.parametrize("geom,deg,ref", [("cube", 2, 3), ("square", 3, 3), ("cube", 1, 3)])
def test_separable_patches_take_the_one_value_per_point_apply(gpu, geom, deg, ref):
    """Boxes are tensor products of 1-D maps: the operator stores one value per
    Gauss point, and its sum-factorised diagonal equals the oracle's
    assembled diagonal."""
    src = (lambda x, p: 1.0 + 0 * x[0]) if geom == "cube" else C.sin_source
    a0 = (lambda x, p: 1.0 + x[0] + 2.0 * x[1] * x[-1])
    case = C.poisson_case(geom, deg, ref, source=src, a0=a0)
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    ctx = M.build_context(case.problem, case.params)
    op = MF.MatrixFreeOperator(ctx, case.params)
    assert op.separable == 1 and op.rational == 0
    v = np.random.default_rng(7).normal(size=K.shape[0])
    y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
    assert np.abs(y - K @ v).max() / np.abs(K @ v).max() < 1e-12
    d = op.diagonal().cpu().numpy()
    np.testing.assert_allclose(d, K.diagonal(), rtol=1e-12)
    d32 = MF.MatrixFreeOperator(ctx, case.params, precision=_lib.IGA_FP32).diagonal().cpu().numpy()
    np.testing.assert_allclose(d32, K.diagonal(), rtol=5e-6)


def test_bf16_coefficient_apply_and_refinement(gpu):
    """BF16 storage of the GRF coefficient (FP32 arithmetic): K v equals the
    oracle's operator assembled on the BF16-rounded coefficient to FP32
    accuracy, differs from the exact one by the BF16 rounding only, and FP64
    iterative refinement on the FP64 operator reaches 1e-10."""
    import scipy.sparse.linalg as spla

    from jaxiga_amd import precision as PR
    from oracle import forms as OF

    case, ctx, a = _grf_cube(3)
    params = {"a0": jx.PointField(a)}
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    a16 = _bf16_round(case.params["a0"].values)
    case16 = C.Case(case.name, case.kind, case.space, case.ospace, case.problem, case.odirichlet,
                    case.oneumann, case.source, {"a0": OF.PointField(a16)})
    _, pat16, data16, _ = _oracle_system(case16)
    K16 = OA.to_scipy(pat16, data16)
    v = np.random.default_rng(9).normal(size=K.shape[0])
    op16 = MF.MatrixFreeOperator(ctx, params, precision=_lib.IGA_BF16)
    y16 = op16.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
    scale = np.abs(K16 @ v).max()
    err16 = np.abs(y16 - K16 @ v).max() / scale
    err_exact = np.abs(y16 - K @ v).max() / scale
    print(f"[mf-bf16] vs rounded-coefficient operator {err16:.2e}, vs exact {err_exact:.2e}", flush=True)
    assert err16 < 5e-6
    assert 1e-5 < err_exact < 1e-2
    op64 = MF.MatrixFreeOperator(ctx, params)
    dinv = MF.inverse_guard(op64.diagonal())
    rhs = torch.as_tensor(b, device="cuda")
    res = PR.mf_mixed_solve(op64, op16, rhs, dinv, tol=1e-10, inner_tol=1e-3)
    x = res.x.cpu().numpy()
    assert np.linalg.norm(b - K @ x) <= 1e-10 * np.linalg.norm(b) * 1.01
    x_ref = spla.spsolve(K.tocsc(), b)
    assert np.abs(x - x_ref).max() <= 1e-8 * np.abs(x_ref).max()


def test_apply_rejects_a_workspace_that_was_not_set_up(gpu):
    case = C.poisson_case("cube", 2, 2, source=lambda x, p: 1.0 + 0 * x[0])
    ctx = M.build_context(case.problem, case.params)
    op = MF.MatrixFreeOperator(ctx, case.params)
    x = torch.zeros(ctx.n_free, dtype=torch.float64, device="cuda")
    from jaxiga_amd.runtime import lib, stream

    # an address no workspace starts at (allocations are 512-byte aligned)
    other = op.work.data_ptr() + 8
    rc = lib().iga_mf_apply(ctx.space.device().struct, other, op.nbytes - 8, ctx.n_free,
                            _lib.ptr(ctx.free_d), _lib.ptr(ctx.f2f_d), _lib.ptr(x), _lib.ptr(x), stream())
    assert rc == _lib.IGA_ERR_ARG
    rc = lib().iga_mf_apply(ctx.space.device().struct, _lib.ptr(op.work), op.nbytes - 256, ctx.n_free,
                            _lib.ptr(ctx.free_d), _lib.ptr(ctx.f2f_d), _lib.ptr(x), _lib.ptr(x), stream())
    assert rc == _lib.IGA_ERR_SHAPE


def _elasticity_cube(deg=2, ref=2, perturb=True):
    """Unit cube, 3-D linear elasticity, clamped at x0 = 0, interior weights
    perturbed by up to +-30 % (rational), E = 2, nu = 0.3."""
    import dataclasses

    from oracle import space as OS

    P, O = C.patches("cube", deg, ref)
    if perturb:
        rng = np.random.default_rng(9)
        w = O.weights * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, size=O.weights.shape))
        P = dataclasses.replace(P, weights=w)
        O = dataclasses.replace(O, weights=w.copy())
    V = jx.FunctionSpace(P, vec=3)
    clamp = lambda x: x[:, 0] < 1e-12  # noqa: E731
    body = lambda x, p: torch.stack([0 * x[0], 0 * x[0], -1.0 + 0 * x[0]]) if isinstance(x, torch.Tensor) \
        else np.stack([0 * x[0], 0 * x[0], -1.0 + 0 * x[0]])  # noqa: E731
    prob = jx.LinearElasticity(V, dirichlet=[jx.DirichletBC(0.0, where=clamp)], source=body)
    return C.Case(f"elasticity-cube-p{deg}-r{ref}", "elasticity", V, OS.function_space(O, vec=3), prob,
                  [C.OF.DirichletBC(0.0, clamp)], [], body, {"E": 2.0, "nu": 0.3}, "stress")


@pytest.mark.parametrize("make", [lambda: C.elasticity_case(3, 2, "strain"),
                                  lambda: C.elasticity_case(2, 3, "stress"),
                                  lambda: _elasticity_cube(2, 2),
                                  lambda: _elasticity_cube(1, 2, perturb=False)])
def test_matrix_free_elasticity_apply_matches_oracle(gpu, make):
    """K v of the sum-factorised linear-elasticity apply (rational annulus
    p=3 and p=2, rational and polynomial cubes) against the oracle's
    assembled operator: FP64 to 1e-12, FP32 to 5e-6."""
    case = make()
    print(f"[mf-elasticity] {case.name}", flush=True)
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    v = np.random.default_rng(11).normal(size=K.shape[0])
    ref_y = K @ v
    ctx = M.build_context(case.problem, case.params)
    for prec, tol in ((_lib.IGA_FP64, 1e-12), (_lib.IGA_FP32, 5e-6)):
        op = MF.MatrixFreeOperator(ctx, case.params, precision=prec)
        assert op.elastic and not op.separable
        y = op.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
        err = np.abs(y - ref_y).max() / np.abs(ref_y).max()
        print(f"[mf-elasticity] precision {prec}: K v rel err {err:.2e}", flush=True)
        assert err < tol


def test_matrix_free_elasticity_solve_matches_assembled(gpu):
    """jx.solve(..., matrix_free=True) on the quarter-annulus elasticity
    problem: the same Jacobi-CG iterates as the assembled operator (iteration
    counts within 1) and the solution within 1e-8 of the oracle's direct
    solution."""
    import scipy.sparse.linalg as spla

    case = C.elasticity_case(3, 2, "strain")
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    x_ref = spla.spsolve(K.tocsc(), b)
    opts = jx.LinearOptions(method="cg", tol=1e-12)
    sol_a = jx.solve(case.problem, params=case.params, linear=opts)
    sol_m = jx.solve(case.problem, params=case.params,
                     linear=jx.LinearOptions(method="cg", tol=1e-12, matrix_free=True))
    assert sol_m.stats["matrix_free"] is True and sol_m.stats["nnz"] is None
    assert abs(sol_m.stats["iterations"] - sol_a.stats["iterations"]) <= 1
    u = sol_m.u.cpu().numpy()[dm.free]
    assert np.abs(u - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
