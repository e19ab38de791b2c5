"""Oracle end-to-end Galerkin solve.

Restates ``src/jaxiga/methods/galerkin.py`` solve_galerkin :275-357 for linear
problems (build context, rhs, CSR assembly, sparse solve, diagnostics, lift)
with ``src/jaxiga/methods/_common.py`` build_context :163-178.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import assembly as A
from oracle import forms as F
from oracle import solver as S
from oracle.space import boundary_basis, gauss_basis, make_dofmap


def element_tangents(kind, space, basis, params, plane="stress"):
    if kind == "poisson":
        return F.poisson_tangents(basis, params, space.vec)
    return F.elasticity_tangents(basis, params, space.dim, plane)


def solve(kind, space, dirichlet=(), neumann=(), source=None, params=None, plane="stress",
          method="cg", tol=1e-10, maxiter=None, quad=None):
    prescribed, values = F.resolve_dirichlet(space, dirichlet, params)
    dofmap = make_dofmap(space.n_dofs, prescribed, values)
    basis = gauss_basis(space, quad)
    Ke = element_tangents(kind, space, basis, params, plane)
    gd = F.gdofs(space)
    pattern = A.csr_pattern(space.elem_dofs, space.n_dofs, space.vec, dofmap.free)
    data = A.assemble_csr(pattern, Ke)

    load = np.zeros(space.n_dofs)
    if source is not None:
        load += A.global_load(space.n_dofs, gd, F.source_load(basis, source, params, space.vec))
    for bc in neumann:
        bset = F.resolve_neumann(space, bc.where)
        bb = boundary_basis(space, bset, quad)
        load += A.global_load(space.n_dofs, F.gdofs(space, bset.elems),
                              F.neumann_load(bb, bc.value, params, space.vec))
    load -= A.lifting(space.n_dofs, gd, Ke, dofmap)
    b = load[dofmap.free]

    K = A.to_scipy(pattern, data)
    iters = None
    if method == "cg":
        u_free, iters = S.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol, maxiter)
    elif method == "scipy_cg":
        u_free, iters = S.scipy_csr_cg(K, b, tol, maxiter)
    else:
        u_free = spla.spsolve(K.tocsc(), b)
    stats = {"method": "galerkin", "linear": method, "matrix_free": False,
             "n_free": len(dofmap.free), "nnz": pattern.nnz, "iterations": iters}
    stats.update(S.linear_diagnostics(lambda v: K @ v, u_free, b, tol))
    return {"u": dofmap.lift(u_free), "u_free": u_free, "b": b, "pattern": pattern, "data": data,
            "dofmap": dofmap, "Ke": Ke, "basis": basis, "stats": stats}
