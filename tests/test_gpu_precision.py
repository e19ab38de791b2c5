"""Mixed precision: FP32 matrix-core assembly and FP32 matrix-free applies,
with FP64 iterative refinement recovering the FP64 oracle's solution.

Tolerances: FP32 stiffness entries within 2e-6 of the oracle relative to the
largest entry; refined solutions at FP64 relative residual <= 1e-10 and within
1e-8 (relative, max norm) of the oracle's FP64 solution."""

import numpy as np
import pytest
import torch

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd import matfree as MF
from jaxiga_amd import methods as M
from jaxiga_amd import precision as PR
from oracle import assembly as OA
from oracle import solver as OSV
from tests import cases as C
from tests.test_gpu_parity import _oracle_system, rational_cube

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: C.elasticity_case(3, 3, "strain"),
                                  lambda: C.poisson_case("annulus", 3, 3),
                                  lambda: rational_cube(2, 3)])
def test_fp32_assembly_entries_and_refined_solution(gpu, make):
    case = make()
    print(f"[mixed-csr] {case.name}", flush=True)
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    u_ref = OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-13)[0]
    ctx = M.build_context(case.problem, case.params)
    v64, rhs, _ = M.assemble(ctx, case.params)
    v32, _, _ = M.assemble(ctx, case.params, load=False, precision=_lib.IGA_FP32)
    assert v32.dtype == torch.float32
    err = np.abs(v32.double().cpu().numpy() - data).max() / np.abs(data).max()
    assert err < 2e-6, err
    dinv = LIN.inverse_diagonal(ctx.assembly, v64)
    res = PR.csr_mixed_solve(ctx.assembly, v64, v32, rhs, dinv, tol=1e-10)
    assert res.residuals[-1] <= 1e-10
    assert res.outer >= 2  # the FP32 operator alone cannot reach 1e-10
    u = res.x.cpu().numpy()
    assert np.abs(u - u_ref).max() / np.abs(u_ref).max() < 1e-8


@pytest.mark.parametrize("make", [lambda: C.elasticity_case(3, 3, "strain"),
                                  lambda: C.poisson_case("annulus", 3, 3)])
def test_fp32_csr_cg_with_reliable_updates(gpu, make):
    # one CG run on the FP32 values, residual replaced by the FP64 one: the
    # stopping test holds for the FP64 operator, with about the FP64 CG's
    # iteration count
    case = make()
    print(f"[mixed-cg] {case.name}", flush=True)
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    u_ref = OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-13)[0]
    ctx = M.build_context(case.problem, case.params)
    v64, rhs, _ = M.assemble(ctx, case.params)
    v32, _, _ = M.assemble(ctx, case.params, load=False, precision=_lib.IGA_FP32)
    dinv = LIN.inverse_diagonal(ctx.assembly, v64)
    ref = LIN.cg_csr(ctx.assembly, v64, rhs, 1e-10, dinv=dinv)
    for every in (16, 64):
        res = PR.csr_mixed_cg(ctx.assembly, v64, v32, rhs, dinv, tol=1e-10, replace_every=every)
        rn, bn = LIN.residual_norms(ctx.assembly, v64, res.x, rhs)
        print(f"[mixed-cg] every {every}: {res.iterations} it (fp64 cg {ref.iterations}), "
              f"true rel. res {rn / bn:.2e}", flush=True)
        assert rn <= 1.0001e-10 * bn
        assert abs(res.relative_residual - rn / bn) <= 1e-3 * rn / bn
        assert res.iterations <= 2 * ref.iterations
        u = res.x.cpu().numpy()
        assert np.abs(u - u_ref).max() / np.abs(u_ref).max() < 1e-8


def test_fp32_matrix_free_refined_solution(gpu):
    case = C.poisson_case("cube", 2, 3, source=lambda x, p: 1.0 + 0 * x[0])
    dm, pat, data, b = _oracle_system(case)
    K = OA.to_scipy(pat, data)
    u_ref = OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-13)[0]
    ctx = M.build_context(case.problem, case.params)
    _, rhs, diag = M.assemble(ctx, case.params, values=False, diagonal=True)
    op64 = MF.MatrixFreeOperator(ctx, case.params)
    op32 = MF.MatrixFreeOperator(ctx, case.params, precision=_lib.IGA_FP32)
    res = PR.mf_mixed_solve(op64, op32, rhs, MF.inverse_guard(diag), tol=1e-10)
    assert res.residuals[-1] <= 1e-10
    u = res.x.cpu().numpy()
    assert np.abs(u - u_ref).max() / np.abs(u_ref).max() < 1e-8
