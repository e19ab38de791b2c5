"""Oracle forms: boundary conditions, element kernels, materials, loads.

Restates
* ``src/jaxiga/forms/problem.py`` -- DirichletBC/Neumann :25-84,
  resolve_dirichlet :179-232, _resolve_dirichlet_dofs :283-302,
  _resolve_neumann_set :305-352;
* ``src/jaxiga/forms/library.py`` -- _strain_displacement :17-37,
  Poisson.element_tangent :54-65, LinearElasticity.element_tangent :174-181,
  _coefficient :228-233;
* ``src/jaxiga/forms/materials.py`` -- plane_stress/plane_strain/isotropic_3d
  :18-72, constitutive :82-88;
* ``src/jaxiga/methods/_common.py`` -- PointField :28-51, element_energy source
  work :304-309, neumann_work :312-318.

User callables are vectorised: a point argument ``x`` is passed as a
``(dim, n)`` array, so code written for one point (``x[0]``, ``x[1]``) works
unchanged on many.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from oracle.space import BoundarySet, boundary_rule, evaluate, expand_dofs


@dataclasses.dataclass(frozen=True)
class DirichletBC:
    value: object
    where: object
    component: int | None = None


@dataclasses.dataclass(frozen=True)
class Neumann:
    value: object
    where: object


@dataclasses.dataclass(frozen=True)
class PointField:
    values: np.ndarray  # (n_e, n_q)


def plane_stress(E, nu):
    return E / (1 - nu**2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]], float)


def plane_strain(E, nu):
    f = E / ((1 + nu) * (1 - 2 * nu))
    return f * np.array([[1 - nu, nu, 0], [nu, 1 - nu, 0], [0, 0, (1 - 2 * nu) / 2]], float)


def isotropic_3d(E, nu):
    f = E / ((1 + nu) * (1 - 2 * nu))
    s = (1 - 2 * nu) / 2
    C = np.zeros((6, 6))
    C[:3, :3] = nu
    np.fill_diagonal(C[:3, :3], 1 - nu)
    C[3, 3] = C[4, 4] = C[5, 5] = s
    return f * C


def constitutive(E, nu, dim, plane="stress"):
    if dim == 3:
        return isotropic_3d(E, nu)
    return plane_stress(E, nu) if plane == "stress" else plane_strain(E, nu)


def strain_displacement(dR):
    """Engineering-Voigt B, ``(..., n_strain, n_l*dim)``, columns (node, comp)."""
    *lead, dim, n_l = dR.shape
    ns = 3 if dim == 2 else 6
    B = np.zeros((*lead, ns, n_l, dim))
    if dim == 2:
        B[..., 0, :, 0] = dR[..., 0, :]
        B[..., 1, :, 1] = dR[..., 1, :]
        B[..., 2, :, 0] = dR[..., 1, :]
        B[..., 2, :, 1] = dR[..., 0, :]
    else:
        for k in range(3):
            B[..., k, :, k] = dR[..., k, :]
        B[..., 3, :, 1] = dR[..., 2, :]
        B[..., 3, :, 2] = dR[..., 1, :]
        B[..., 4, :, 0] = dR[..., 2, :]
        B[..., 4, :, 2] = dR[..., 0, :]
        B[..., 5, :, 0] = dR[..., 1, :]
        B[..., 5, :, 1] = dR[..., 0, :]
    return B.reshape(*lead, ns, n_l * dim)


def coefficient(params, name, default, x):
    """Per-point coefficient ``(n_e, n_q)``; ``x`` is ``(n_e, n_q, dim)``."""
    if params is None or name not in params:
        return np.full(x.shape[:2], float(default))
    v = params[name]
    if isinstance(v, PointField):
        return np.asarray(v.values, float)
    if callable(v):
        flat = x.reshape(-1, x.shape[-1]).T
        return np.broadcast_to(np.asarray(v(flat, params), float), flat.shape[1:]).reshape(x.shape[:2])
    return np.full(x.shape[:2], float(v))


def poisson_tangents(basis, params, vec=1):
    a0 = coefficient(params, "a0", 1.0, basis.x)
    K = np.einsum("eqdi,eqdj->eij", basis.dR * (basis.w * a0)[:, :, None, None], basis.dR,
                  optimize=True)
    if vec == 1:
        return K
    n_e, n_l, _ = K.shape
    out = np.einsum("eij,cd->eicjd", K, np.eye(vec))
    return out.reshape(n_e, n_l * vec, n_l * vec)


def elasticity_tangents(basis, params, dim, plane="stress"):
    C = constitutive(params["E"], params["nu"], dim, plane)
    B = strain_displacement(basis.dR)
    CB = np.einsum("st,eqtj->eqsj", C, B, optimize=True) * basis.w[:, :, None, None]
    return np.einsum("eqsi,eqsj->eij", B, CB, optimize=True)


def _call_points(fn, x, *extra):
    """Evaluate a vectorised user function at ``(n_e, n_q, dim)`` points;
    returns ``(n_e, n_q, vec)``."""
    flat = x.reshape(-1, x.shape[-1]).T
    args = [e.reshape(-1, e.shape[-1]).T for e in extra]
    out = np.asarray(fn(flat, *args), float)
    n = flat.shape[1]
    if out.ndim == 0:
        out = np.full((1, n), float(out))
    elif out.ndim == 1:
        out = out[None, :] if out.shape[0] == n else out[:, None] * np.ones((1, n))
    return out.T.reshape(*x.shape[:2], -1)


def source_load(basis, source, params, vec):
    """``F[(i,c)] = sum_q w f_c R_i`` per element, ``(n_e, n_l*vec)``."""
    f = _call_points(lambda xx: source(xx, params), basis.x)
    f = np.broadcast_to(f, (*basis.x.shape[:2], vec))
    F = np.einsum("eq,eqc,eqi->eic", basis.w, f, basis.R)
    return F.reshape(F.shape[0], -1)


def neumann_load(basis, traction, params, vec):
    t = _call_points(lambda xx, nn: traction(xx, nn, params), basis.x, basis.normal)
    t = np.broadcast_to(t, (*basis.x.shape[:2], vec))
    F = np.einsum("eq,eqc,eqi->eic", basis.w, t, basis.R)
    return F.reshape(F.shape[0], -1)


def _per_point(value, n):
    """Normalise a vectorised callable's result to ``(k, n)``, k = 1 or vec."""
    v = np.asarray(value, float)
    if v.ndim == 0:
        return np.full((1, n), float(v))
    if v.ndim == 1:
        return v[None, :] if v.shape[0] == n else np.repeat(v[:, None], n, 1)
    return v


def resolve_dirichlet(space, bcs, params=None):
    """Sorted prescribed dofs and values, nodal interpolation at control points."""
    entries = {}
    all_boundary = np.unique(np.concatenate([b.dofs for b in space.boundaries.values()]))
    for bc in bcs:
        if isinstance(bc.where, str):
            sdofs = space.boundary(bc.where).dofs
        else:
            mask = np.asarray(bc.where(space.cpts[all_boundary]), bool)
            sdofs = all_boundary[mask]
            if sdofs.size == 0:
                raise ValueError("Dirichlet predicate selected no boundary dofs")
        comps = range(space.vec) if bc.component is None else [bc.component]
        if callable(bc.value):
            vals = _per_point(bc.value(space.cpts[sdofs].T, params), len(sdofs))
        else:
            # a literal: one value, or one per component (forms/problem.py:203-206)
            vals = np.repeat(np.atleast_1d(np.asarray(bc.value, float))[:, None], len(sdofs), 1)
        for k, sd in enumerate(sdofs):
            for c in comps:
                g = int(space.vec * sd + c)
                v = float(vals[c, k] if vals.shape[0] == space.vec else vals[0, k])
                if g in entries and abs(entries[g] - v) > 1e-12:
                    raise ValueError(f"conflicting Dirichlet conditions on dof {g}")
                entries[g] = v
    dofs = np.array(sorted(entries), dtype=np.int64)
    return dofs, np.array([entries[int(d)] for d in dofs], float)


def side_centres(space, elems, sides):
    ref = np.zeros((len(elems), 1, space.dim))
    for i, code in enumerate(sides):
        d, end = divmod(int(code), 2)
        ref[i, 0, d] = 1.0 if end else -1.0
    return evaluate(space, elems, ref, None, sides).x[:, 0, :]


def resolve_neumann(space, where):
    if isinstance(where, str):
        return space.boundary(where)
    elems = np.concatenate([b.elems for b in space.boundaries.values()])
    sides = np.concatenate([b.sides for b in space.boundaries.values()])
    _, uniq = np.unique(np.stack([elems, sides], 1), axis=0, return_index=True)
    keep = np.sort(uniq)
    elems, sides = elems[keep], sides[keep]
    mask = np.asarray(where(side_centres(space, elems, sides)), bool).ravel()
    if not mask.any():
        raise ValueError("Neumann predicate selected no boundary elements")
    ke, ks = elems[mask], sides[mask]
    dofs = np.unique(space.elem_dofs[ke].ravel())
    return BoundarySet(ke, ks, dofs)


def gdofs(space, elems=None):
    ed = space.elem_dofs if elems is None else space.elem_dofs[elems]
    return expand_dofs(ed, space.vec)


__all__ = [
    "DirichletBC", "Neumann", "PointField", "constitutive", "poisson_tangents",
    "elasticity_tangents", "source_load", "neumann_load", "resolve_dirichlet",
    "resolve_neumann", "boundary_rule",
]
