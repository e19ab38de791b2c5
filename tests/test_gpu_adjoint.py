"""Implicit differentiation of the Galerkin solve on the GPU (torch autograd
through ``jx.solve`` and ``linear.sparse_solve``) against the oracle's
reverse pass (oracle/adjoint.py, itself pinned by finite differences) and
against central finite differences of the GPU forward solve, as the
reference's gradient tests do (tests/test_galerkin.py:183-274).

Bars: cotangents within 1e-7 of the oracle's relative to their scale (the
GPU solves with CG to a 1e-13 relative residual, the oracle directly);
gradients within the reference's finite-difference tolerances; exactly one
forward and one adjoint solve per gradient.
"""

import numpy as np
import pytest
import torch

import jaxiga_amd as jx
from jaxiga_amd import adjoint as AD
from jaxiga_amd import linear as LIN
from jaxiga_amd import methods as M
from jaxiga_amd.evaluation import evaluate
from jaxiga_amd.pointset import gauss
from oracle import adjoint as OJ
from oracle import forms as OF
from oracle import geometry as OG
from oracle import pipeline as OP
from oracle import space as OS
from tests.cases import UNIT, everywhere, xp

pytestmark = pytest.mark.gpu

TIGHT = jx.LinearOptions(method="cg", preconditioner="jacobi", tol=1e-13)
QUAD = [[0, 0], [1.2, 0.1], [1.0, 1.0], [-0.1, 0.9]]


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def lin_g(x, p=None):
    return 0.3 + x[0] - 0.5 * x[1]


def amp_source(x, p):
    return p["amp"] * (1.0 + x[0] * x[1])


def _poisson_pair(deg=2, ref=2):
    V = jx.FunctionSpace(jx.primitives.quadrilateral(QUAD).elevate(deg).refine(ref))
    VO = OS.function_space(OG.quadrilateral(QUAD).elevate(deg).refine(ref))
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(lin_g, where=everywhere)], source=amp_source)
    return V, VO, prob


def test_field_points_match_oracle(gpu):
    for vec, (geom, deg, ref) in ((1, ("annulus", 3, 2)), (2, ("square", 2, 2))):
        P = (jx.primitives.quarter_annulus(1.0, 2.0) if geom == "annulus"
             else jx.primitives.quadrilateral(UNIT)).elevate(deg).refine(ref)
        O = (OG.quarter_annulus(1.0, 2.0) if geom == "annulus" else OG.quadrilateral(UNIT)).elevate(deg).refine(ref)
        V, VO = jx.FunctionSpace(P, vec=vec), OS.function_space(O, vec=vec)
        rng = np.random.default_rng(vec)
        u0, u1 = rng.standard_normal(V.n_dofs), rng.standard_normal(V.n_dofs)
        basis = evaluate(V, gauss(V))
        val, grad = AD.field_points(basis, torch.as_tensor(u0, device=gpu), torch.as_tensor(u1, device=gpu), vec)
        ob = OS.gauss_basis(VO)
        for f, u in enumerate((u0, u1)):
            ul = u[OF.gdofs(VO)].reshape(VO.n_elems, -1, vec)
            ov = np.einsum("eqi,eic->eqc", ob.R, ul)
            og = np.einsum("eqdi,eic->eqcd", ob.dR, ul)
            print(f"[field] vec={vec} field {f}: val {rel(val[f].cpu(), ov):.1e} grad {rel(grad[f].cpu(), og):.1e}",
                  flush=True)
            assert rel(val[f].cpu(), ov) < 1e-12 and rel(grad[f].cpu(), og) < 1e-12
        sol = jx.Solution(space=V, u=torch.as_tensor(u0, device=gpu))
        assert rel(sol.at(basis=basis).cpu(), val[0].cpu()) == 0.0
        assert rel(sol.grad(basis=basis).cpu(), grad[0].cpu()) == 0.0


def test_sparse_solve_reverse_rule_matches_oracle(gpu):
    V, VO, prob = _poisson_pair()
    a0 = 1.0 + np.random.default_rng(3).random((V.n_elems, 9))
    params = {"a0": jx.PointField(torch.as_tensor(a0, device=gpu)), "amp": 1.5}
    ctx = M.build_context(prob, params)
    vals, b, _ = M.assemble(ctx, params)
    vals.requires_grad_()
    b.requires_grad_()
    c = torch.as_tensor(np.random.default_rng(4).standard_normal(ctx.n_free), device=gpu)
    LIN.reset_solve_counter()
    u, info = LIN.sparse_solve(TIGHT, ctx.assembly, vals, b)
    (c * u).sum().backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}
    out = OP.solve("poisson", VO, [OF.DirichletBC(lin_g, everywhere)], source=amp_source,
                   params={"a0": OF.PointField(a0), "amp": 1.5}, method="direct")
    data_bar, lam = OJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c.cpu().numpy())
    print(f"[vjp] data_bar {rel(vals.grad.cpu(), data_bar):.1e} b_bar {rel(b.grad.cpu(), lam):.1e}", flush=True)
    assert rel(vals.grad.cpu(), data_bar) < 1e-7
    assert rel(b.grad.cpu(), lam) < 1e-7


def test_poisson_gradients_match_oracle_and_finite_differences(gpu):
    V, VO, prob = _poisson_pair()
    rng = np.random.default_rng(5)
    a0 = 1.0 + rng.random((V.n_elems, 9))
    c = rng.standard_normal(V.n_dofs)
    ct = torch.as_tensor(c, device=gpu)
    a0_t = torch.as_tensor(a0, device=gpu).requires_grad_()
    amp_t = torch.tensor(1.5, dtype=torch.float64, device=gpu, requires_grad=True)
    LIN.reset_solve_counter()
    sol = jx.solve(prob, params={"a0": jx.PointField(a0_t), "amp": amp_t}, linear=TIGHT)
    (ct * sol.u).sum().backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}

    out = OP.solve("poisson", VO, [OF.DirichletBC(lin_g, everywhere)], source=amp_source,
                   params={"a0": OF.PointField(a0), "amp": 1.5}, method="direct")
    dm = out["dofmap"]
    db, lam = OJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c[dm.free])
    gd = OF.gdofs(VO)
    Kbar, load_bar = OJ.element_cotangent(out["pattern"], db, lam, gd, dm)
    oc = OJ.point_cotangents("poisson", out["basis"], Kbar, load_bar, gd, 1)
    x = out["basis"].x
    d_amp = float((oc["source"][..., 0] * (1.0 + x[..., 0] * x[..., 1])).sum())
    print(f"[grad] a0 points {rel(a0_t.grad.cpu(), oc['a0']):.1e}; amp {amp_t.grad.item():.10e} "
          f"oracle {d_amp:.10e}", flush=True)
    assert rel(a0_t.grad.cpu(), oc["a0"]) < 1e-7
    assert abs(amp_t.grad.item() - d_amp) <= 1e-7 * abs(d_amp)

    def loss(amp):
        with torch.no_grad():
            return float((ct * jx.solve(prob, params={"a0": jx.PointField(torch.as_tensor(a0, device=gpu)),
                                                      "amp": amp}, linear=TIGHT).u).sum())

    h = 1e-5
    np.testing.assert_allclose(amp_t.grad.item(), (loss(1.5 + h) - loss(1.5 - h)) / (2 * h), rtol=1e-6)


def test_vjp_uses_exactly_one_adjoint_solve(gpu):
    """Scalar and callable coefficients; loss sum(u^2), as the reference's
    test_vjp_uses_exactly_one_adjoint_solve (tests/test_galerkin.py:264-274)."""
    V = jx.FunctionSpace(jx.primitives.quadrilateral(UNIT).elevate(2).refine(2))
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=everywhere)],
                      source=lambda x, p: 1.0 + 0.0 * x[0])

    def loss(a0, grad=True):
        with torch.set_grad_enabled(grad):
            return (jx.solve(prob, params={"a0": a0}).u ** 2).sum()

    a0 = torch.tensor(1.3, dtype=torch.float64, device=gpu, requires_grad=True)
    LIN.reset_solve_counter()
    loss(a0).backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}
    h = 1e-6
    fd = (loss(1.3 + h, False) - loss(1.3 - h, False)).item() / (2 * h)
    print(f"[one-adjoint] d/da0 {a0.grad.item():.10e} fd {fd:.10e}", flush=True)
    np.testing.assert_allclose(a0.grad.item(), fd, rtol=1e-6)
    # a coefficient field through a callable of a parameter
    k = torch.tensor(0.7, dtype=torch.float64, device=gpu, requires_grad=True)
    field = lambda kk: (lambda x, p: kk * (1.0 + xp(x[0]).cos(x[0]) * x[1]))  # noqa: E731
    LIN.reset_solve_counter()
    loss(field(k)).backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}
    fd = (loss(field(0.7 + h), False) - loss(field(0.7 - h), False)).item() / (2 * h)
    np.testing.assert_allclose(k.grad.item(), fd, rtol=1e-6)


def test_matrix_free_vjp_costs_one_adjoint_solve(gpu):
    """The reference's tests/test_advanced.py:272-283: the matrix-free solve
    (CG at 1e-13) differentiated with one adjoint solve, against a central
    difference; and the gradient in a0 at every Gauss point equals the
    assembled path's."""
    opts = jx.LinearOptions(method="cg", matrix_free=True, tol=1e-13)
    V = jx.FunctionSpace(jx.primitives.quadrilateral(QUAD).elevate(2).refine(2))
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(lin_g, where=everywhere)], source=amp_source)

    def loss(a0, linear, grad=True):
        with torch.set_grad_enabled(grad):
            return (jx.solve(prob, params={"a0": a0, "amp": 1.5}, linear=linear).u ** 2).sum()

    a0 = torch.tensor(1.0, dtype=torch.float64, device=gpu, requires_grad=True)
    LIN.reset_solve_counter()
    loss(a0, opts).backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}
    h = 1e-6
    fd = (loss(1.0 + h, opts, False) - loss(1.0 - h, opts, False)).item() / (2 * h)
    print(f"[matfree] d/da0 {a0.grad.item():.12e} fd {fd:.12e}", flush=True)
    assert abs(a0.grad.item() - fd) / abs(fd) < 1e-7
    a = 1.0 + torch.rand((V.n_elems, 9), dtype=torch.float64, device=gpu,
                         generator=torch.Generator(device=gpu).manual_seed(8))
    grads = []
    for linear in (opts, TIGHT):
        at = a.clone().requires_grad_()
        loss(jx.PointField(at), linear).backward()
        grads.append(at.grad.cpu().numpy())
    print(f"[matfree] point gradients vs assembled: {rel(grads[0], grads[1]):.1e}", flush=True)
    assert rel(grads[0], grads[1]) < 1e-9


def _elastic_traction(x, n, p):
    m = xp(x)
    return m.stack([p["t"] * (1.0 + 0.0 * x[0]), 0.3 * p["t"] * x[1]])


def test_elasticity_gradients_match_oracle_and_finite_differences(gpu):
    V = jx.FunctionSpace(jx.primitives.rectangle(0, 0, 2, 1).elevate(2).refine(2), vec=2)
    VO = OS.function_space(OG.rectangle(0, 0, 2, 1).elevate(2).refine(2), vec=2)
    shear = lambda x, p=None: 0.01 * x[1]  # noqa: E731
    prob = jx.LinearElasticity(V, plane="strain",
                               dirichlet=[jx.DirichletBC(shear, where="left", component=0),
                                          jx.DirichletBC(0.0, where="left", component=1)],
                               neumann=[jx.Neumann(_elastic_traction, where="right")])
    E0, nu0, t0 = 1e3, 0.3, 2.0
    rng = np.random.default_rng(6)
    c = rng.standard_normal(V.n_dofs)
    ct = torch.as_tensor(c, device=gpu)
    th = torch.tensor([E0, nu0, t0], dtype=torch.float64, device=gpu, requires_grad=True)
    LIN.reset_solve_counter()
    sol = jx.solve(prob, params={"E": th[0], "nu": th[1], "t": th[2]}, linear=TIGHT)
    (ct * sol.u).sum().backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}
    g = th.grad.cpu().numpy()

    out = OP.solve("elasticity", VO, [OF.DirichletBC(shear, "left", component=0),
                                      OF.DirichletBC(0.0, "left", component=1)],
                   [OF.Neumann(_elastic_traction, "right")], params={"E": E0, "nu": nu0, "t": t0},
                   plane="strain", method="direct")
    dm = out["dofmap"]
    db, lam = OJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c[dm.free])
    gd = OF.gdofs(VO)
    Kbar, load_bar = OJ.element_cotangent(out["pattern"], db, lam, gd, dm)
    oc = OJ.point_cotangents("elasticity", out["basis"], Kbar, load_bar, gd, 2)
    Cf = lambda e_, n_: OF.constitutive(e_, n_, 2, "strain")  # noqa: E731
    hn = 1e-6
    bset = OF.resolve_neumann(VO, "right")
    bb = OS.boundary_basis(VO, bset)
    tb = OJ.boundary_cotangent(bb, load_bar, OF.gdofs(VO, bset.elems), 2)
    ref = np.array([(oc["C"] * Cf(1.0, nu0)).sum(),
                    (oc["C"] * (Cf(E0, nu0 + hn) - Cf(E0, nu0 - hn)) / (2 * hn)).sum(),
                    (tb * np.stack([np.ones(bb.x.shape[:2]), 0.3 * bb.x[..., 1]], -1)).sum()])
    print(f"[elastic] grad {g} oracle {ref}", flush=True)
    np.testing.assert_allclose(g, ref, rtol=1e-7)

    def loss(E, nu, t):
        with torch.no_grad():
            return float((ct * jx.solve(prob, params={"E": E, "nu": nu, "t": t}, linear=TIGHT).u).sum())

    # the reference's steps and tolerance (tests/test_galerkin.py:183-195)
    for i, h in ((0, 1.0), (1, 1e-6), (2, 1e-5)):
        tp, tm = [E0, nu0, t0], [E0, nu0, t0]
        tp[i] += h
        tm[i] -= h
        np.testing.assert_allclose(g[i], (loss(*tp) - loss(*tm)) / (2 * h), rtol=1e-5)


def test_shape_derivative_wrt_control_points_reference_case(gpu):
    """The reference's tests/test_galerkin.py:233-262: compliance sum(u^2) of
    a p=2, 2x2 unit-square Poisson solve (u = 0 on the boundary, f = 1,
    a0 = 1) differentiated in the control points, rebuilt on the moved
    points; finite differences at flat indices 5 and 11 (rtol 1e-4, atol
    1e-8), one forward and one adjoint solve."""
    V = jx.FunctionSpace(jx.primitives.quadrilateral(UNIT).elevate(2).refine(1))
    base = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=everywhere)], source=lambda x, p: 1.0 + 0 * x[0])

    def compliance(cpts):
        problem = base.with_space(V.replace(cpts=cpts))
        return (jx.solve(problem, params={"a0": 1.0}, linear=TIGHT).u ** 2).sum()

    c0 = torch.tensor(V.cpts, dtype=torch.float64, requires_grad=True)
    LIN.reset_solve_counter()
    compliance(c0).backward()
    assert LIN.SOLVE_COUNTER == {"forward": 1, "adjoint": 1}, LIN.SOLVE_COUNTER
    g = c0.grad.reshape(-1).numpy()
    assert np.any(np.abs(g) > 1e-12)
    flat = torch.tensor(V.cpts, dtype=torch.float64).reshape(-1)
    for i in (5, 11):
        h = 1e-6
        with torch.no_grad():
            fp = flat.clone()
            fp[i] += h
            fm = flat.clone()
            fm[i] -= h
            fd = float(compliance(fp.reshape(V.cpts.shape)) - compliance(fm.reshape(V.cpts.shape))) / (2 * h)
        print(f"[shape] index {i}: adjoint {g[i]:.9e} finite difference {fd:.9e}", flush=True)
        np.testing.assert_allclose(g[i], fd, rtol=1e-4, atol=1e-8)


def _shape_oracle(kind):
    """Matched GPU and oracle problems for the shape cotangent: distorted quad
    Poisson (a0 and source callables of x, constant non-zero Dirichlet
    value) and plane-strain elasticity on the rational quarter annulus with a
    body force of x."""
    if kind == "poisson":
        P = jx.primitives.quadrilateral(QUAD).elevate(2).refine(2)
        O = OG.quadrilateral(QUAD).elevate(2).refine(2)
        V, VO = jx.FunctionSpace(P), OS.function_space(O)
        a0 = lambda x, p: 1.0 + x[0] ** 2 + 0.5 * x[1]  # noqa: E731
        src = lambda x, p: 1.0 + x[0] * x[1]  # noqa: E731
        prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.3, where=everywhere)], source=src)
        return V, VO, prob, [OF.DirichletBC(0.3, everywhere)], src, {"a0": a0}, "stress"
    P = jx.primitives.quarter_annulus(1.0, 2.0).elevate(2).refine(1)
    O = OG.quarter_annulus(1.0, 2.0).elevate(2).refine(1)
    V, VO = jx.FunctionSpace(P, vec=2), OS.function_space(O, vec=2)

    def src(x, p):
        m = xp(x)
        return m.stack([0.2 + x[1], -1.0 + 0.3 * x[0] * x[1]])

    prob = jx.LinearElasticity(V, plane="strain", dirichlet=[jx.DirichletBC(0.0, where="inner")], source=src)
    return V, VO, prob, [OF.DirichletBC(0.0, "inner")], src, {"E": 1.0, "nu": 0.3}, "strain"


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_shape_cotangent_matches_oracle(gpu, kind):
    """d(c.u)/dX against the oracle's restated shape cotangent (pinned by
    finite differences of its own solves): 1e-7 relative to the cotangent's
    scale (the oracle differentiates the callables by central differences)."""
    V, VO, prob, odir, src, params, plane = _shape_oracle(kind)
    rng = np.random.default_rng(8)
    c = torch.as_tensor(rng.standard_normal(V.n_dofs), device="cuda")
    X = torch.tensor(V.cpts, dtype=torch.float64, requires_grad=True)
    sol = jx.solve(prob.with_space(V.replace(cpts=X)), params=params, linear=TIGHT)
    (c * sol.u).sum().backward()
    out = OP.solve(kind, VO, odir, source=src, params=params, plane=plane, method="direct")
    from oracle import assembly as OA

    K = OA.to_scipy(out["pattern"], out["data"])
    import scipy.sparse.linalg as spla

    cc = c.cpu().numpy()
    lam = np.zeros(VO.n_dofs)
    lam[out["dofmap"].free] = spla.spsolve(K.tocsc(), cc[out["dofmap"].free])
    ref = OJ.shape_cotangent(kind, VO, OS.gauss_basis(VO), out["u"], lam, params, src, plane=plane)
    err = rel(X.grad.numpy(), ref)
    print(f"[shape] {kind}: rel. difference to the oracle {err:.2e}", flush=True)
    assert err < 1e-7


def test_shape_derivative_rejects_what_it_does_not_cover(gpu):
    V = jx.FunctionSpace(jx.primitives.quadrilateral(UNIT).elevate(2).refine(1))
    X = torch.tensor(V.cpts, dtype=torch.float64, requires_grad=True)
    prob = jx.Poisson(V.replace(cpts=X), dirichlet=[jx.DirichletBC(lambda x, p: x[0], where=everywhere)],
                      source=lambda x, p: 1.0 + 0 * x[0])
    with pytest.raises(NotImplementedError, match="constant"):
        jx.solve(prob, params={"a0": 1.0}, linear=TIGHT)
