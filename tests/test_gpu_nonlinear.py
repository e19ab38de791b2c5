"""The nonlinear Galerkin path on the GPU (Hyperelasticity, Newton with the
tangent re-assembled into the same CSR pattern every step) against the
oracle (oracle/nonlinear.py, pinned by finite differences of its energy):
residual within 1e-12 and tangent CSR values within 1e-12 of their scale at
a random displacement; the Newton solution within 1e-9 with the same step
count; the reference's own nonlinear case (tests/test_methods.py:106-127)
converging quadratically in at most 10 steps."""

import numpy as np
import pytest
import torch

import jaxiga_amd as jx
from jaxiga_amd import methods as M
from jaxiga_amd import nonlinear as NLG
from oracle import forms as OF
from oracle import geometry as OG
from oracle import nonlinear as ONL
from oracle import space as OS
from tests.cases import UNIT, xp

pytestmark = pytest.mark.gpu

SQUARE_PARAMS = {"mu": 384.6, "lam": 576.9}


def _square(deg=2, ref=1):
    V = jx.FunctionSpace(jx.primitives.quadrilateral(UNIT).elevate(deg).refine(ref), vec=2)
    VO = OS.function_space(OG.quadrilateral(UNIT).elevate(deg).refine(ref), vec=2)
    prob = jx.Hyperelasticity(V, dirichlet=[jx.DirichletBC([0.0, 0.0], where="bottom"),
                                            jx.DirichletBC([0.0, 0.2], where="top")])
    osys = ONL.NeoHookeanSystem(VO, [OF.DirichletBC([0.0, 0.0], "bottom"), OF.DirichletBC([0.0, 0.2], "top")],
                                params=SQUARE_PARAMS)
    return prob, osys, SQUARE_PARAMS


def _loaded(deg=2, ref=2):
    """reference tests/test_implicit_diff.py:75-88: left edge clamped, a
    traction on the right edge, plus a body force of x here."""
    params = {"E": 200.0, "nu": 0.3, "load": 5.0}

    def traction(x, n, p):
        m = xp(x)
        return m.stack([p["load"] + 0 * x[0], 0.3 * p["load"] + 0 * x[0]])

    def body(x, p):
        m = xp(x)
        return m.stack([0.5 * x[1], -1.0 + 0 * x[0]])

    V = jx.FunctionSpace(jx.primitives.quadrilateral(UNIT).elevate(deg).refine(ref), vec=2)
    VO = OS.function_space(OG.quadrilateral(UNIT).elevate(deg).refine(ref), vec=2)
    prob = jx.Hyperelasticity(V, dirichlet=[jx.DirichletBC([0.0, 0.0], where="left")],
                              neumann=[jx.Neumann(traction, where="right")], source=body)
    osys = ONL.NeoHookeanSystem(VO, [OF.DirichletBC([0.0, 0.0], "left")],
                                [OF.Neumann(traction, "right")], body=body, params=params)
    return prob, osys, params


@pytest.mark.parametrize("make", [_square, _loaded])
def test_residual_and_tangent_match_oracle(gpu, make):
    prob, osys, params = make()
    ctx = M.build_context(prob, params)
    np.testing.assert_array_equal(ctx.dofmap.free, osys.dofmap.free)
    op = NLG.NeoHookeanOperator(ctx, params)
    rng = np.random.default_rng(2)
    u = 0.02 * rng.standard_normal(ctx.n_free)
    r = op.residual(torch.as_tensor(u, device="cuda")).cpu().numpy()
    r_ref = osys.residual(u)
    assert np.abs(r - r_ref).max() <= 1e-12 * np.abs(r_ref).max()
    vals = op.tangent(torch.as_tensor(u, device="cuda")).cpu().numpy()
    np.testing.assert_array_equal(ctx.assembly.column_indices, osys.pattern.column_indices)
    k_ref = osys.tangent(u)
    assert np.abs(vals - k_ref).max() <= 1e-12 * np.abs(k_ref).max()


@pytest.mark.parametrize("make", [_square, _loaded])
def test_newton_matches_oracle(gpu, make):
    prob, osys, params = make()
    sol = jx.solve(prob, params=params, linear=jx.LinearOptions(method="dense"))
    st = sol.stats
    u_ref, st_ref = ONL.newton(osys)
    print(f"[newton] steps {st['iterations']} (oracle {st_ref['iterations']}), history {st['history']}",
          flush=True)
    assert st["converged"] and st_ref["converged"]
    assert st["iterations"] == st_ref["iterations"]
    u = sol.u.cpu().numpy()
    ref = osys.full(u_ref)
    assert np.abs(u - ref).max() <= 1e-9 * np.abs(ref).max()


def test_reference_nonlinear_case_converges_quadratically(gpu):
    """reference tests/test_methods.py:106-127 with the default linear
    options (SuperLU, as the reference picks for a problem not declared SPD)."""
    prob, _, params = _square()
    sol = jx.solve(prob, params=params)
    h = sol.stats["history"]
    assert sol.stats["converged"] and sol.stats["iterations"] <= 10
    assert h[-1] < 1e-9 and h[-1] < h[-2] ** 1.5
    top = prob.space.boundary("top").dofs
    np.testing.assert_allclose(sol.u.reshape(-1, 2)[top, 1].cpu().numpy(), 0.2)


def test_newton_with_cg_on_the_tangent(gpu):
    prob, osys, params = _square(2, 2)
    sol = jx.solve(prob, params=params, linear=jx.LinearOptions(method="cg", tol=1e-12))
    u_ref, _ = ONL.newton(osys)
    assert sol.stats["converged"]
    ref = osys.full(u_ref)
    assert np.abs(sol.u.cpu().numpy() - ref).max() <= 1e-8 * np.abs(ref).max()
