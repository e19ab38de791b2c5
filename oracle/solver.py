"""Oracle solvers, diagnostics and error norms.

Restates
* ``src/jaxiga/solvers/linear.py`` -- LinearOptions validation :102-127;
  _scipy_csr_cg :261-321 (SciPy CG on native CSR, Jacobi via the diagonal
  with a zero guard, target ``0.5*tol*||b||``, up to three true-residual
  correction solves sharing ``maxiter`` = 10 n by default); the accelerator
  semantics of ``jax.scipy.sparse.linalg.cg`` used by _iterative_csr :324-351
  (x0 = 0, stop when ``||r||_2 <= tol*||b||_2``, Jacobi inverse diagonal with
  zero guard);
* ``src/jaxiga/solvers/diagnostics.py`` -- linear_diagnostics :8-19;
* ``src/jaxiga/post/norms.py`` -- errornorm :59-111, norm :42-56,
  integrate :35-39;
* the textbook Jacobi-PCG written in ``tests/test_advanced.py:250-264``.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle.space import expand_dofs, gauss_basis


def inverse_diagonal(diag):
    inv = np.ones_like(diag)
    np.divide(1.0, diag, out=inv, where=diag != 0)
    return inv


def scipy_csr_cg(A, b, tol=1e-10, maxiter=None, preconditioner="jacobi"):
    """Returns ``(u, iterations)``."""
    M = None
    if preconditioner == "jacobi":
        inv = inverse_diagonal(A.diagonal())
        M = spla.LinearOperator(A.shape, matvec=lambda v: inv * v, dtype=A.dtype)
    b = np.asarray(b, float)
    u = np.zeros_like(b)
    res = b.copy()
    nrm = np.linalg.norm(res)
    target = 0.5 * tol * nrm
    left = maxiter if maxiter is not None else 10 * len(b)
    total = 0
    for attempt in range(4):
        if nrm <= target or left <= 0:
            break
        used = [0]

        def tick(_):
            used[0] += 1

        corr, info = spla.cg(A, res, M=M, rtol=min(0.5, 0.5 * target / nrm), atol=0.0,
                             maxiter=left, callback=tick)
        left -= used[0]
        total += used[0]
        cand = u + corr
        actual = b - A @ cand
        an = np.linalg.norm(actual)
        if attempt > 0 and an >= nrm:
            break
        u, res, nrm = cand, actual, an
        if info != 0 or used[0] == 0 or not np.isfinite(nrm):
            break
    return u, total


def jacobi_pcg(matvec, b, diag, tol=1e-10, maxiter=None):
    """Accelerator-path CG semantics; returns ``(u, iterations)``."""
    b = np.asarray(b, float)
    n = len(b)
    maxiter = 10 * n if maxiter is None else maxiter
    inv = inverse_diagonal(np.asarray(diag, float))
    u = np.zeros(n)
    r = b.copy()
    z = inv * r
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    it = 0
    while np.linalg.norm(r) > tol * bn and it < maxiter:
        Ap = matvec(p)
        alpha = rz / (p @ Ap)
        u += alpha * p
        r -= alpha * Ap
        z = inv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return u, it


def linear_diagnostics(matvec, u, b, tol):
    norm = float(np.linalg.norm(matvec(u) - b))
    scale = float(np.linalg.norm(b))
    rel = norm / (scale if scale > 0 else 1.0)
    finite = bool(np.isfinite(norm) and np.all(np.isfinite(u)))
    return {"residual": norm, "relative_residual": rel, "converged": finite and norm <= tol * scale,
            "finite": finite}


def _local(space, u):
    idx = expand_dofs(space.elem_dofs, space.vec)
    return u[idx].reshape(space.n_elems, space.n_local, space.vec)


def field_at(space, u, basis):
    return np.einsum("eqn,enc->eqc", basis.R, _local(space, u))


def grad_at(space, u, basis):
    return np.einsum("eqdn,enc->eqcd", basis.dR, _local(space, u))


def _user(fn, x, out_shape):
    flat = x.reshape(-1, x.shape[-1]).T
    v = np.asarray(fn(flat), float)
    if v.ndim >= 1 and v.shape[0] == flat.shape[1]:
        v = v.reshape(flat.shape[1], -1)
    else:
        v = np.moveaxis(v.reshape(-1, flat.shape[1]), -1, 0)
    return v.reshape(out_shape)


def errornorm(space, u, exact, norm="L2", exact_grad=None, quad=None, relative=True):
    basis = gauss_basis(space, quad)
    w = basis.w
    uh = field_at(space, u, basis)
    if norm == "L2":
        ue = _user(exact, basis.x, uh.shape)
        num = np.sum(w * np.sum((uh - ue) ** 2, -1))
        den = np.sum(w * np.sum(ue**2, -1))
    elif norm == "H1":
        gh = grad_at(space, u, basis)
        ge = _user(exact_grad, basis.x, gh.shape)
        num = np.sum(w * np.sum((gh - ge) ** 2, axis=(-1, -2)))
        den = np.sum(w * np.sum(ge**2, axis=(-1, -2)))
    else:
        raise ValueError(f"unknown norm {norm!r}")
    return float(np.sqrt(num / den) if relative else np.sqrt(num))


def l2_norm(space, u, quad=None):
    basis = gauss_basis(space, quad)
    uh = field_at(space, u, basis)
    return float(np.sqrt(np.sum(basis.w * np.sum(uh**2, -1))))


def integrate(space, fn, quad=None):
    basis = gauss_basis(space, quad)
    flat = basis.x.reshape(-1, basis.x.shape[-1]).T
    v = np.asarray(fn(flat), float).reshape(basis.w.shape)
    return float(np.sum(basis.w * v))
