"""Oracle of the nonlinear Galerkin path -- TEST INFRASTRUCTURE ONLY.

Restates, in NumPy/SciPy FP64, the compressible neo-Hookean solid of the
reference (``src/jaxiga/forms/library.py:206-225``: psi = mu/2 (tr F^T F - d)
- mu ln J + lam/2 (ln J)^2, F = I + grad u, Lame parameters from E and nu by
``forms/materials.py:75-79``) and Newton's method with backtracking on the
residual norm (``src/jaxiga/solvers/newton.py:27-145``) as the Galerkin
driver calls it (``src/jaxiga/methods/galerkin.py:337-355``). The reference
obtains the residual and tangent by JAX autodiff of the energy; here they are
written out (first Piola stress and its derivative) and pinned by finite
differences of the energy (``tests/test_oracle_nonlinear.py``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import assembly as A
from oracle import forms as F
from oracle.space import boundary_basis, gauss_basis, make_dofmap

MAX_BACKTRACKS = 20


def lame(E, nu):
    """``(lam, mu)`` (``forms/materials.py:75-79``)."""
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def material(params):
    if "mu" in params:
        return float(params["mu"]), float(params["lam"])
    lam, mu = lame(float(params["E"]), float(params["nu"]))
    return mu, lam


def _deformation(space, basis, u):
    ul = np.asarray(u, float).reshape(-1, space.dim)[space.elem_dofs]  # (n_e, n_l, dim)
    return np.eye(space.dim) + np.einsum("eqja,eai->eqij", basis.dR, ul)


def energy(space, basis, u, mu, lam):
    """Internal energy ``sum_q w psi(F)``."""
    Fd = _deformation(space, basis, u)
    J = np.linalg.det(Fd)
    I1 = np.einsum("eqij,eqij->eq", Fd, Fd)
    psi = mu / 2 * (I1 - space.dim) - mu * np.log(J) + lam / 2 * np.log(J) ** 2
    return float((basis.w * psi).sum())


def element_residual_tangent(space, basis, u, mu, lam, body=None):
    """Per element ``r (n_e, n_l*dim)`` = internal force minus body force and
    ``K (n_e, n_l*dim, n_l*dim)``, rows/columns ``a*dim + i``."""
    dim = space.dim
    Fd = _deformation(space, basis, u)
    Fi = np.linalg.inv(Fd)  # Fi[j][i] = (F^-T)_ij
    J = np.linalg.det(Fd)
    c = lam * np.log(J) - mu
    P = mu * Fd + c[..., None, None] * np.swapaxes(Fi, -1, -2)
    dR = basis.dR  # (n_e, n_q, dim, n_l)
    r = np.einsum("eq,eqij,eqja->eai", basis.w, P, dR)
    if body is not None:
        r -= np.einsum("eq,eqi,eqa->eai", basis.w, body, basis.R)
    g = np.einsum("eqji,eqja->eqai", Fi, dR)  # g_a = F^-T grad R_a
    dd = np.einsum("eqja,eqjb->eqab", dR, dR)
    K = (mu * np.einsum("eq,eqab,ik->eaibk", basis.w, dd, np.eye(dim))
         + lam * np.einsum("eq,eqai,eqbk->eaibk", basis.w, g, g)
         - np.einsum("eq,eq,eqak,eqbi->eaibk", basis.w, c, g, g))
    n_e, n_l = r.shape[:2]
    return r.reshape(n_e, n_l * dim), K.reshape(n_e, n_l * dim, n_l * dim)


class NeoHookeanSystem:
    """Residual and tangent on the free dofs of a neo-Hookean problem with
    fixed (resolved) Dirichlet values."""

    def __init__(self, space, dirichlet=(), neumann=(), body=None, params=None):
        self.space = space
        self.mu, self.lam = material(params)
        presc, vals = F.resolve_dirichlet(space, dirichlet, params)
        self.dofmap = make_dofmap(space.n_dofs, presc, vals)
        self.basis = gauss_basis(space)
        self.body = None if body is None else np.broadcast_to(
            F._call_points(lambda xx: body(xx, params), self.basis.x), (*self.basis.x.shape[:2], space.dim))
        self.gd = F.gdofs(space)
        self.pattern = A.csr_pattern(space.elem_dofs, space.n_dofs, space.vec, self.dofmap.free)
        ext = np.zeros(space.n_dofs)
        for bc in neumann:
            bset = F.resolve_neumann(space, bc.where)
            bb = boundary_basis(space, bset)
            ext += A.global_load(space.n_dofs, F.gdofs(space, bset.elems),
                                 F.neumann_load(bb, bc.value, params, space.vec))
        self.ext = ext

    def full(self, u_free):
        return self.dofmap.lift(u_free)

    def residual(self, u_free):
        r, _ = element_residual_tangent(self.space, self.basis, self.full(u_free), self.mu, self.lam,
                                        self.body)
        g = A.global_load(self.space.n_dofs, self.gd, r) - self.ext
        return g[self.dofmap.free]

    def tangent(self, u_free):
        _, K = element_residual_tangent(self.space, self.basis, self.full(u_free), self.mu, self.lam,
                                        self.body)
        return A.assemble_csr(self.pattern, K)


def newton(system, u0=None, tol=1e-8, max_steps=25, line_search=True):
    """Newton with backtracking (halving up to 20 times until the residual
    norm decreases); ``(u_free, stats)`` with the reference's stats keys."""
    n = len(system.dofmap.free)
    u = np.zeros(n) if u0 is None else np.asarray(u0, float).copy()

    def norm(v):
        return float(np.linalg.norm(system.residual(v)))

    r_norm = norm(u)
    hist = [r_norm]
    step, alpha = 0, 1.0
    while np.isfinite(r_norm) and r_norm > tol and step < max_steps and alpha > 0:
        r = system.residual(u)
        K = A.to_scipy(system.pattern, system.tangent(u))
        du = spla.spsolve(K.tocsc(), -r)
        alpha = 1.0
        if line_search:
            trial, k = norm(u + du), 0
            while (not np.isfinite(trial) or trial >= r_norm) and k < MAX_BACKTRACKS:
                alpha *= 0.5
                trial = norm(u + alpha * du)
                k += 1
            alpha = alpha if (np.isfinite(trial) and trial < r_norm) else 0.0
        if alpha > 0:
            u = u + alpha * du
            r_norm = norm(u)
        step += 1
        hist.append(r_norm)
    converged = r_norm <= tol
    stalled = alpha <= 0 and not converged
    r_norm = norm(u)
    finite = bool(np.isfinite(r_norm) and np.all(np.isfinite(u)))
    return u, {"iterations": step, "residual": r_norm, "converged": finite and r_norm <= tol,
               "finite": finite, "iteration_limit": step >= max_steps and not converged,
               "stalled": stalled, "history": hist}
