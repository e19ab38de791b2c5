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
point_cotangents the transpose of the element kernels
                  (``src/jaxiga/forms/library.py:40-65,146-181``,
                  ``src/jaxiga/methods/galerkin.py:37-119``) and of the source and
                  Neumann loads: cotangents of ``a0`` and the source at the interior
                  Gauss points, of ``C``, and of tractions at boundary points

Pinned by central finite differences of the oracle's own direct solve
(``tests/test_oracle_adjoint.py``).
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
