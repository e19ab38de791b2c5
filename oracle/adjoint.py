"""Oracle reverse pass of the Galerkin solve -- TEST INFRASTRUCTURE ONLY.

The reference differentiates ``jx.solve`` with ``jax.grad``: JAX autodiff
through the element kernels and the scatter-add, and a custom reverse rule for
the sparse solve. This module restates both in NumPy/SciPy FP64, the general
way (through the element-matrix cotangent), as the checker of the product's
reverse pass:

csr_solve_vjp     ``_csr_sparse_solve_bwd`` (``src/jaxiga/solvers/linear.py:376-386``):
                  one adjoint solve on the same symmetric CSR,
                  ``data_bar = -lam[rows] u[cols]``, ``b_bar = lam``
element_cotangent the transpose of the scatter-add (``assemble_csr``,
                  ``src/jaxiga/methods/galerkin.py:122-148``) and of the lifting
                  ``-K u_g`` (``src/jaxiga/methods/_common.py:280-375``): ``Kbar_e``
shape_cotangent  the cotangent of the control points (shape derivative): the
                  reference differentiates the whole basis evaluation
                  (``src/jaxiga/space/evaluation.py:95-172``: x = R X, J = dR X,
                  w = wref |det J|, grad = J^-1 grad_xi) and the element kernels
                  with JAX; restated as the first variation of the point
                  functional under a motion of the map
point_cotangents the transpose of the element kernels
                  (``src/jaxiga/forms/library.py:40-65,146-181``,
                  ``src/jaxiga/methods/galerkin.py:37-119``) and of the source and
                  Neumann loads: cotangents of ``a0`` and the source at the interior
                  Gauss points, of ``C``, and of tractions at boundary points

Pinned by central finite differences of the oracle's own direct solve
(``tests/test_oracle_adjoint.py``), the shape cotangent by finite
differences of solves on perturbed control points.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import assembly as A
from oracle import forms as F


def csr_solve_vjp(pattern, data, u_free, u_bar):
    """``(data_bar, b_bar)`` of ``u = K(data)^-1 b`` for the cotangent ``u_bar``."""
    K = A.to_scipy(pattern, data)
    lam = spla.spsolve(K.tocsc(), np.asarray(u_bar, float))
    data_bar = -lam[pattern.row_indices] * np.asarray(u_free)[pattern.column_indices]
    return data_bar, lam


def element_cotangent(pattern, data_bar, b_bar, gdofs, dofmap):
    """``Kbar_e (n_e, n_loc, n_loc)``: ``data_bar`` gathered through the
    scatter map (transpose of the scatter-add), plus the lifting's part
    ``load = ... - K u_g`` with ``b = load[free]``: ``-load_bar_i u_g,j``."""
    n_e = gdofs.shape[0]
    Kbar = np.append(np.asarray(data_bar, float), 0.0)[pattern.scatter].reshape(n_e, gdofs.shape[1], -1)
    load_bar = np.zeros(dofmap.n_dofs)
    load_bar[dofmap.free] = b_bar
    ug = np.zeros(dofmap.n_dofs)
    ug[dofmap.prescribed] = dofmap.values
    Kbar = Kbar - np.einsum("ei,ej->eij", load_bar[gdofs], ug[gdofs])
    return Kbar, load_bar


def point_cotangents(kind, basis, Kbar, load_bar, gdofs, vec, dim=None):
    """Cotangents at the interior Gauss points: ``a0`` ``(n_e, n_q)`` (Poisson)
    or ``C`` ``(n_s, n_s)`` (elasticity), and the source ``(n_e, n_q, vec)``."""
    out = {}
    if kind == "poisson":
        n_e, n_q, d, n_l = basis.dR.shape
        Kb = Kbar.reshape(n_e, n_l, vec, n_l, vec)
        Ks = np.einsum("eicjc->eij", Kb)  # K_e = K_scalar (x) I_vec
        out["a0"] = basis.w * np.einsum("eqdi,eij,eqdj->eq", basis.dR, Ks, basis.dR, optimize=True)
    else:
        B = F.strain_displacement(basis.dR)  # (n_e, n_q, n_s, n_l*dim)
        out["C"] = np.einsum("eq,eqsi,eij,eqtj->st", basis.w, B, Kbar, B, optimize=True)
    lb = load_bar[gdofs].reshape(gdofs.shape[0], -1, vec)  # (n_e, n_l, vec)
    out["source"] = basis.w[..., None] * np.einsum("eqi,eic->eqc", basis.R, lb)
    return out


def boundary_cotangent(bbasis, load_bar, gdofs_b, vec):
    """Cotangent of the traction at boundary points ``(n_e_b, n_q_b, vec)``."""
    lb = load_bar[gdofs_b].reshape(gdofs_b.shape[0], -1, vec)
    return bbasis.w[..., None] * np.einsum("eqi,eic->eqc", bbasis.R, lb)


def _stress_tensor(sv, dim):
    """Symmetric tensor ``(..., dim, dim)`` from a Voigt vector (xx, yy, xy) /
    (xx, yy, zz, yz, xz, xy)."""
    T = np.zeros((*sv.shape[:-1], dim, dim))
    if dim == 2:
        T[..., 0, 0], T[..., 1, 1] = sv[..., 0], sv[..., 1]
        T[..., 0, 1] = T[..., 1, 0] = sv[..., 2]
    else:
        for k in range(3):
            T[..., k, k] = sv[..., k]
        T[..., 1, 2] = T[..., 2, 1] = sv[..., 3]
        T[..., 0, 2] = T[..., 2, 0] = sv[..., 4]
        T[..., 0, 1] = T[..., 1, 0] = sv[..., 5]
    return T


def _voigt(g):
    """Engineering Voigt strain from ``g[..., i, j] = d u_i / d x_j``."""
    if g.shape[-1] == 2:
        return np.stack([g[..., 0, 0], g[..., 1, 1], g[..., 0, 1] + g[..., 1, 0]], -1)
    return np.stack([g[..., 0, 0], g[..., 1, 1], g[..., 2, 2], g[..., 1, 2] + g[..., 2, 1],
                     g[..., 0, 2] + g[..., 2, 0], g[..., 0, 1] + g[..., 1, 0]], -1)


def _point_values(fn, x, params, k):
    """A point callable at ``x (n_e, n_q, dim)``: ``(n_e, n_q, k)``."""
    return np.broadcast_to(F._call_points(lambda xx: fn(xx, params), x), (*x.shape[:2], k))


def _x_gradient(fn, x, params, k, h=1e-6):
    """``d fn / d x`` at the points by central differences, ``(n_e, n_q, k, dim)``."""
    out = np.empty((*x.shape[:2], k, x.shape[-1]))
    for d in range(x.shape[-1]):
        e = np.zeros(x.shape[-1])
        e[d] = h
        out[..., d] = (_point_values(fn, x + e, params, k) - _point_values(fn, x - e, params, k)) / (2 * h)
    return out


def shape_cotangent(kind, space, basis, u, lam, params=None, source=None, plane="stress"):
    """``d/dX [lam^T (F(X) - K(X) u)]`` over the control points ``(n_cp, dim)``:
    the shape cotangent of a solve whose adjoint is ``lam`` (zero on
    prescribed dofs) and whose full solution is ``u``. Under a motion
    ``dX_k`` of control point k every point moves by ``R_k dX_k``, so
    ``d w = w div V``, ``d grad(phi) = -grad(V)^T grad(phi)`` and the point
    functional ``w (f.lam - eps(lam):sigma(u))`` varies by
    ``sum_q P : grad(R_k) + R_k xbar`` with
    ``P_cj = w [(f.lam - eps(lam):sigma(u)) delta_cj
               + sum_i (d_c lam_i sigma_ij(u) + d_c u_i sigma_ij(lam))]``
    (``sigma = a grad`` for Poisson) and ``xbar = w d/dx (f.lam - eps:sigma)``
    through the explicit ``x`` of callable coefficients and sources
    (central differences here)."""
    vec, dim = space.vec, space.dim
    R, dR, x, w = basis.R, basis.dR, basis.x, basis.w
    ed = space.elem_dofs
    ul = np.asarray(u, float).reshape(-1, vec)[ed]  # (n_e, n_l, vec)
    ll = np.asarray(lam, float).reshape(-1, vec)[ed]
    gu = np.einsum("eqda,eai->eqid", dR, ul)  # d u_i / d x_d
    gl = np.einsum("eqda,eai->eqid", dR, ll)
    lq = np.einsum("eqa,eai->eqi", R, ll)
    P = np.zeros((*x.shape[:2], dim, dim))
    xbar = np.zeros(x.shape)
    f = np.zeros((*x.shape[:2], vec)) if source is None else _point_values(source, x, params, vec)
    fl = (f * lq).sum(-1)
    if kind == "poisson":
        a = F.coefficient(params, "a0", 1.0, x)
        inner = (gu[:, :, 0, :] * gl[:, :, 0, :]).sum(-1)
        phi = fl - a * inner
        P += phi[..., None, None] * np.eye(dim)
        P += a[..., None, None] * (gl[:, :, 0, :, None] * gu[:, :, 0, None, :] +
                                   gu[:, :, 0, :, None] * gl[:, :, 0, None, :])
        a0 = None if params is None else params.get("a0")
        if callable(a0) and not isinstance(a0, F.PointField):
            xbar -= inner[..., None] * _x_gradient(a0, x, params, 1)[:, :, 0, :]
    else:
        C = F.constitutive(params["E"], params["nu"], dim, plane)
        su = _stress_tensor(np.einsum("st,eqt->eqs", C, _voigt(gu)), dim)
        sl = _stress_tensor(np.einsum("st,eqt->eqs", C, _voigt(gl)), dim)
        energy = (_voigt(gl) * np.einsum("st,eqt->eqs", C, _voigt(gu))).sum(-1)
        P += (fl - energy)[..., None, None] * np.eye(dim)
        P += np.einsum("eqic,eqij->eqcj", gl, su) + np.einsum("eqic,eqij->eqcj", gu, sl)
    if source is not None:
        xbar += np.einsum("eqi,eqid->eqd", lq, _x_gradient(source, x, params, vec))
    P *= w[..., None, None]
    xbar *= w[..., None]
    Gl = np.einsum("eqcj,eqja->eac", P, dR) + np.einsum("eqa,eqc->eac", R, xbar)
    G = np.zeros((space.n_scalar_basis, dim))
    np.add.at(G, ed.reshape(-1), Gl.reshape(-1, dim))
    return G
