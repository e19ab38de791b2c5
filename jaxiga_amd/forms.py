"""Problems, boundary conditions and materials (the element-kernel plug-in seam).

Mirrors ``forms/problem.py`` (``DirichletBC`` :25-48, ``Neumann`` :75-84,
``Problem`` :87-141, ``resolve_dirichlet`` :179-232,
``_resolve_dirichlet_dofs`` :283-302, ``_resolve_neumann_set`` :305-352),
``forms/library.py`` (``Poisson`` :40-65, ``LinearElasticity`` :146-203,
``_coefficient`` :228-233), ``forms/materials.py`` (:18-112) and
``methods/_common.py`` ``PointField`` :28-51.

A built-in problem names its element kernel (``kernel``) instead of being
differentiated: the Galerkin driver hands the kernel's parameters to the
fused HIP assembly. Setup-stage callables (Dirichlet predicates and values)
run on host NumPy like the reference's; point callables (sources, callable
coefficients, tractions) receive ``x`` as a ``(dim, n)`` float64 tensor on the
GPU -- single-point code written as ``x[0]``, ``x[1]`` vectorises unchanged --
and must return a tensor (or a Python scalar).
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Sequence

import numpy as np

from jaxiga_amd.space import BoundarySet


@dataclasses.dataclass(frozen=True)
class DirichletBC:
    value: float | Callable
    where: str | Callable
    component: int | None = None


@dataclasses.dataclass(frozen=True)
class Neumann:
    value: Callable
    where: str | Callable


@dataclasses.dataclass(frozen=True)
class PointField:
    """Per-Gauss-point data ``(n_elems, n_q)`` passed through ``params``."""

    values: object

    @property
    def shape(self):
        return tuple(self.values.shape)


# ---------------------------------------------------------------------------
# materials (engineering Voigt notation)
# ---------------------------------------------------------------------------


def _is_tensor(*xs):
    return any(type(x).__module__.split(".")[0] == "torch" for x in xs)


def _matrix(scale, rows):
    """``scale * rows`` as a float64 torch tensor when an entry is a torch
    tensor (so gradients reach ``E`` and ``nu``), else as a NumPy array."""
    flat = [x for r in rows for x in r]
    if not _is_tensor(scale, *flat):
        return float(scale) * np.array(rows, float)
    import torch

    ref = next(x for x in (scale, *flat) if _is_tensor(x))
    t = lambda x: torch.as_tensor(x, dtype=torch.float64, device=ref.device)  # noqa: E731
    return t(scale) * torch.stack([torch.stack([t(x) for x in r]) for r in rows])


class materials:
    @staticmethod
    def plane_stress(E, nu):
        return _matrix(E / (1 - nu**2), [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])

    @staticmethod
    def plane_strain(E, nu):
        return _matrix(E / ((1 + nu) * (1 - 2 * nu)),
                       [[1 - nu, nu, 0], [nu, 1 - nu, 0], [0, 0, (1 - 2 * nu) / 2]])

    @staticmethod
    def isotropic_3d(E, nu):
        d, s = 1 - nu, (1 - 2 * nu) / 2
        rows = [[d, nu, nu, 0, 0, 0], [nu, d, nu, 0, 0, 0], [nu, nu, d, 0, 0, 0],
                [0, 0, 0, s, 0, 0], [0, 0, 0, 0, s, 0], [0, 0, 0, 0, 0, s]]
        return _matrix(E / ((1 + nu) * (1 - 2 * nu)), rows)

    @staticmethod
    def lame(E, nu):
        return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))

    @staticmethod
    def constitutive(E, nu, dim, plane="stress"):
        if dim == 3:
            return materials.isotropic_3d(E, nu)
        if dim == 2:
            return materials.plane_stress(E, nu) if plane == "stress" else materials.plane_strain(E, nu)
        raise ValueError(f"elasticity needs dim 2 or 3, got {dim}")


# ---------------------------------------------------------------------------
# problems
# ---------------------------------------------------------------------------


class Problem:
    """Base class: a space, boundary conditions and a source."""

    is_linear = False
    is_positive_definite = False
    needs_hessian = False
    kernel = None

    def __init__(self, space, *, dirichlet: Sequence[DirichletBC] = (),
                 neumann: Sequence[Neumann] = (), source: Callable | None = None,
                 quadrature: int | None = None):
        self.space = space
        self.dirichlet = tuple(dirichlet)
        self.neumann = tuple(neumann)
        self.source = source
        self.quadrature = quadrature
        for bc in self.dirichlet:
            if bc.component is not None and not (0 <= bc.component < space.vec):
                raise ValueError(f"DirichletBC component {bc.component} is out of range for a "
                                 f"space with vec={space.vec}")
        self._dirichlet_dofs = tuple((bc, _resolve_dirichlet_dofs(space, bc)) for bc in self.dirichlet)
        self._neumann_sets = tuple((bc, _resolve_neumann_set(space, bc.where)) for bc in self.neumann)

    def energy(self, grad_u, u, x, params):
        raise NotImplementedError(
            f"{type(self).__name__} defines no energy(); use a built-in problem "
            "(Poisson, LinearElasticity)")

    def with_space(self, space) -> "Problem":
        """The same problem on another space of the same connectivity (the
        geometry changes, the physics and the resolved boundary dofs do not),
        as ``Problem.with_space`` of the reference (``forms/problem.py:167-175``)
        for shape derivatives."""
        clone = object.__new__(type(self))
        clone.__dict__.update(self.__dict__)
        clone.space = space
        return clone

    def resolve_dirichlet(self, params=None):
        """Sorted prescribed global dofs and their nodal values at control points."""
        space = self.space
        entries = {}
        for bc_idx, (bc, sdofs) in enumerate(self._dirichlet_dofs):
            comps = range(space.vec) if bc.component is None else [bc.component]
            if callable(bc.value):
                vals = _per_point(bc.value(space.cpts[sdofs].T, params), len(sdofs))
            else:
                # one value, or one per component (forms/problem.py:203-206)
                vals = np.repeat(np.atleast_1d(np.asarray(bc.value, float))[:, None], len(sdofs), 1)
            k = space.vec
            for c in comps:
                row = vals[c] if vals.shape[0] == space.vec else vals[0]
                g = k * np.asarray(sdofs, np.int64) + c
                for gi, v in zip(g.tolist(), row.tolist()):
                    if gi in entries and abs(entries[gi] - v) > 1e-12:
                        raise ValueError(f"conflicting Dirichlet conditions on dof {gi}: boundary "
                                         f"condition {bc_idx} prescribes a different value")
                    entries[gi] = v
        if not entries:
            return np.zeros(0, dtype=np.int64), np.zeros(0)
        dofs = np.array(sorted(entries), dtype=np.int64)
        return dofs, np.array([entries[int(d)] for d in dofs], float)

    def resolve_neumann(self):
        return list(self._neumann_sets)


class Poisson(Problem):
    """``-div(a0 grad u) = f`` from ``psi = 1/2 a0 |grad u|^2``; ``params['a0']``
    may be a scalar, a callable ``a0(x, params)`` or a :class:`PointField`."""

    is_linear = True
    is_positive_definite = True
    kernel = "poisson"


class LinearElasticity(Problem):
    """Small-strain elasticity ``psi = 1/2 eps : C : eps``; ``params`` holds
    ``E`` and ``nu``; ``plane`` selects plane stress or strain in 2-D."""

    is_linear = True
    is_positive_definite = True
    kernel = "elasticity"

    def __init__(self, space, *, plane: str = "stress", **kwargs):
        if plane not in ("stress", "strain"):
            raise ValueError(f"plane must be 'stress' or 'strain', got {plane!r}")
        if space.vec != space.dim:
            raise ValueError(f"elasticity needs vec == dim; got vec={space.vec}, dim={space.dim}. "
                             f"Build the space with FunctionSpace(patch, vec={space.dim}).")
        self.plane = plane
        super().__init__(space, **kwargs)

    def C(self, params):
        return materials.constitutive(params["E"], params["nu"], self.space.dim, self.plane)


class Hyperelasticity(Problem):
    """Compressible neo-Hookean solid ``psi = mu/2 (tr F^T F - d) - mu ln J
    + lam/2 (ln J)^2``, ``F = I + grad u`` (``forms/library.py:206-225``),
    solved by Newton's method; ``params`` holds ``mu`` and ``lam`` or ``E``
    and ``nu`` (converted by :meth:`materials.lame`)."""

    is_linear = False
    is_positive_definite = False
    kernel = "neohooke"

    def __init__(self, space, **kwargs):
        if space.vec != space.dim:
            raise ValueError(f"hyperelasticity needs vec == dim; got vec={space.vec}, dim={space.dim}. "
                             f"Build the space with FunctionSpace(patch, vec={space.dim}).")
        super().__init__(space, **kwargs)

    @staticmethod
    def lame_parameters(params):
        """``(mu, lam)`` from ``params``."""
        if "mu" in params:
            return float(params["mu"]), float(params["lam"])
        lam, mu = materials.lame(float(params["E"]), float(params["nu"]))
        return mu, lam


# ---------------------------------------------------------------------------
# setup-stage resolution
# ---------------------------------------------------------------------------


def _per_point(value, n):
    v = np.asarray(value, float)
    if v.ndim == 0:
        return np.full((1, n), float(v))
    if v.ndim == 1:
        return v[None, :] if v.shape[0] == n else np.repeat(v[:, None], n, 1)
    return v


def _resolve_dirichlet_dofs(space, bc):
    if isinstance(bc.where, str):
        return np.asarray(space.boundary(bc.where).dofs)
    all_boundary = np.unique(np.concatenate([np.asarray(b.dofs) for b in space.boundaries.values()]))
    mask = np.asarray(bc.where(space.cpts[all_boundary]), dtype=bool)
    if mask.shape != all_boundary.shape:
        raise ValueError(f"Dirichlet predicate returned shape {mask.shape}, expected "
                         f"{all_boundary.shape} (one bool per boundary control point)")
    selected = all_boundary[mask]
    if selected.size == 0:
        raise ValueError("Dirichlet predicate selected no boundary dofs")
    return selected


def _resolve_neumann_set(space, where):
    if isinstance(where, str):
        return space.boundary(where)
    from jaxiga_amd.evaluation import evaluate
    from jaxiga_amd.pointset import PointSet

    elems = np.concatenate([np.asarray(b.elems) for b in space.boundaries.values()])
    sides = np.concatenate([np.asarray(b.sides) for b in space.boundaries.values()])
    _, unique = np.unique(np.stack([elems, sides], axis=1), axis=0, return_index=True)
    keep = np.sort(unique)
    elems, sides = elems[keep], sides[keep]
    ref = np.zeros((len(elems), 1, space.dim))
    for i, code in enumerate(sides):
        d, end = divmod(int(code), 2)
        ref[i, 0, d] = 1.0 if end else -1.0
    ps = PointSet(elems=elems, ref=ref, sides=sides, tag="side-centres")
    centres = evaluate(space, ps, outputs=("x",)).x[:, 0, :].cpu().numpy()
    mask = np.asarray(where(centres), dtype=bool).ravel()
    if mask.shape != elems.shape:
        raise ValueError(f"Neumann predicate returned shape {mask.shape}, expected {elems.shape}")
    if not mask.any():
        raise ValueError("Neumann predicate selected no boundary elements")
    ke, ks = elems[mask], sides[mask]
    return BoundarySet(elems=ke, sides=ks, dofs=np.unique(space.elem_dofs[ke].ravel()))
