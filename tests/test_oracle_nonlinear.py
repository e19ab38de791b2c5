"""Pins the oracle's neo-Hookean residual and tangent (oracle/nonlinear.py)
by central finite differences of its energy and residual, and its Newton
iteration by quadratic convergence on the reference's nonlinear case
(tests/test_methods.py:106-127: unit square, p=2, refine 1, bottom clamped,
top displaced by 0.2)."""

import numpy as np

from oracle import assembly as A
from oracle import forms as F
from oracle import geometry as G
from oracle import nonlinear as NL
from oracle import space as S


def _case():
    V = S.function_space(G.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(2).refine(1), vec=2)
    bcs = [F.DirichletBC([0.0, 0.0], "bottom"), F.DirichletBC([0.0, 0.2], "top")]
    return V, NL.NeoHookeanSystem(V, bcs, params={"mu": 384.6, "lam": 576.9})


def test_residual_and_tangent_match_finite_differences():
    V, sysm = _case()
    rng = np.random.default_rng(0)
    u = 0.02 * rng.standard_normal(V.n_dofs)
    body = 0.5 + rng.random((*sysm.basis.x.shape[:2], 2))
    r, K = NL.element_residual_tangent(V, sysm.basis, u, sysm.mu, sysm.lam, body)
    rg = A.global_load(V.n_dofs, sysm.gd, r)
    load = A.global_load(V.n_dofs, sysm.gd, np.einsum("eq,eqi,eqa->eai", sysm.basis.w, body,
                                                       sysm.basis.R).reshape(V.n_elems, -1))
    h = 1e-6
    for trial in range(3):
        du = rng.standard_normal(V.n_dofs)
        fd = (NL.energy(V, sysm.basis, u + h * du, sysm.mu, sysm.lam)
              - NL.energy(V, sysm.basis, u - h * du, sysm.mu, sysm.lam)) / (2 * h)
        np.testing.assert_allclose((rg + load) @ du, fd, rtol=1e-6)
        rp, _ = NL.element_residual_tangent(V, sysm.basis, u + h * du, sysm.mu, sysm.lam, body)
        rm, _ = NL.element_residual_tangent(V, sysm.basis, u - h * du, sysm.mu, sysm.lam, body)
        fdr = (A.global_load(V.n_dofs, sysm.gd, rp) - A.global_load(V.n_dofs, sysm.gd, rm)) / (2 * h)
        Kd = A.global_load(V.n_dofs, sysm.gd, np.einsum("eij,ej->ei", K, du[sysm.gd]))
        np.testing.assert_allclose(Kd, fdr, rtol=1e-5, atol=1e-7 * np.abs(fdr).max())


def test_newton_converges_quadratically():
    V, sysm = _case()
    u, st = NL.newton(sysm)
    h = st["history"]
    assert st["converged"] and st["iterations"] <= 10
    assert h[-1] < 1e-9 and h[-1] < h[-2] ** 1.5
    # the top edge carries the prescribed displacement
    uf = sysm.full(u).reshape(-1, 2)
    top = V.boundary("top").dofs
    np.testing.assert_allclose(uf[top, 1], 0.2)
