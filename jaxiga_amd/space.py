"""Function space over one tensor-product NURBS patch, DOF maps, and the
device-resident patch tables the HIP kernels read.

Mirrors ``space/function_space.py`` of the reference: ``FunctionSpace``
:182-275 (extraction, cpts, wgts, elem_vertex, boundaries, elem_dofs, vec,
degree), ``expand_dofs`` :300-312, ``_tensor_extraction`` :315-323,
``_patch_connectivity`` :326-369 (element numbering and IEN, first direction
fastest), ``_classify_boundaries`` :490-546, ``DofMap``/``make_dofmap``
:554-603; and ``geometry/multipatch.py`` ``side_indices`` :99-131.

Layout on the device (one copy per space, built lazily):
* per direction ``d``: the 1-D Bezier extraction operators
  ``ext[d] (n_elem_d, p+1, p+1)``, the first supported basis index of every
  span, and the span bounds -- the tensor operator of an element is the
  Kronecker product of these, so nothing per element is stored;
* control points ``(n_cp, dim)`` unweighted and weights ``(n_cp,)``, first
  direction fastest;
* per direction the inclusive 1-D neighbour range of every basis function,
  from which the CSR sparsity is generated on the device.

The 1-D extraction operators are obtained as the exact change of basis from
the spline space to its Bezier (fully broken) refinement: row ``a`` of
``C_e`` lists the Bernstein coefficients of the ``a``-th function supported
on span ``e``.
"""

from __future__ import annotations

import dataclasses
import functools
import hashlib
import weakref

import numpy as np

from jaxiga_amd.geometry import Patch, change_of_basis, side_code, side_name


def element_spans(knots, p):
    """``lo, hi, first`` of every non-empty span (``bernstein.py:261-284``)."""
    t = np.asarray(knots, float)
    idx = np.nonzero(t[1:] > t[:-1])[0]
    return t[idx], t[idx + 1], (idx - p).astype(np.int64)


def extraction_1d(knots, p):
    """``(n_elem, p+1, p+1)`` with ``N_e = C_e @ B`` (Bernstein on [-1, 1])."""
    t = np.asarray(knots, float)
    lo, hi, first = element_spans(t, p)
    vals = np.unique(t)
    bez = np.concatenate([np.full(p + 1, vals[0]), np.repeat(vals[1:-1], p), np.full(p + 1, vals[-1])])
    T = change_of_basis(t, p, bez, p)  # (n_bez, n)
    ops = np.empty((len(lo), p + 1, p + 1))
    for e in range(len(lo)):
        ops[e] = T[e * p : e * p + p + 1, first[e] : first[e] + p + 1].T
    # exact zeros where the change of basis leaves rounding dust
    ops[np.abs(ops) < 1e-15] = 0.0
    return ops


def expand_dofs(scalar_dofs, vec: int):
    scalar_dofs = np.asarray(scalar_dofs)
    if vec == 1:
        return scalar_dofs
    out = vec * scalar_dofs[..., :, None] + np.arange(vec)
    return out.reshape(*scalar_dofs.shape[:-1], -1)


def _block_indices(ranges, strides):
    """Flat indices of a tensor block, first listed range fastest."""
    idx = np.zeros(1, dtype=np.int64)
    for r, s in zip(ranges, strides):
        idx = (np.asarray(r, dtype=np.int64)[:, None] * s + idx[None, :]).reshape(-1)
    return idx


def side_indices(patch: Patch, side: str, offset: int = 0) -> np.ndarray:
    direction = "uvw".index(side[0])
    end = int(side[1:])
    n = patch.n_cp_per_dir
    if not 0 <= offset < n[direction]:
        raise ValueError(f"offset {offset} is outside the {n[direction]} control points")
    strides = np.concatenate([[1], np.cumprod(n[:-1])]).astype(np.int64)
    ranges = [np.arange(k) for k in n]
    ranges[direction] = np.array([offset if end == 0 else n[direction] - 1 - offset])
    return _block_indices(ranges, strides)


@dataclasses.dataclass(frozen=True)
class BoundarySet:
    elems: np.ndarray
    sides: np.ndarray
    dofs: np.ndarray


class FunctionSpace:
    """Global NURBS basis over one tensor-product patch."""

    def __init__(self, patch: Patch, vec: int = 1):
        if isinstance(patch, (list, tuple)):
            if len(patch) != 1:
                raise NotImplementedError("multipatch spaces are outside the supported hot path")
            patch = patch[0]
        self.patch = patch
        self.vec = int(vec)
        self.dim = patch.dim
        self.degree = tuple(patch.degree)
        self.cell_type = "tensor"
        knots = patch.knot_arrays()
        spans = [element_spans(knots[d], self.degree[d]) for d in range(self.dim)]
        self.extraction_ops = [extraction_1d(knots[d], self.degree[d]) for d in range(self.dim)]
        self.n_elem_per_dir = tuple(len(s[0]) for s in spans)
        self._spans = spans
        self.n_elems = int(np.prod(self.n_elem_per_dir))
        self.n_scalar_basis = patch.n_cp
        self.cpts = np.asarray(patch.ctrl_pts, float)
        self.wgts = np.asarray(patch.weights, float)

        ne = self.n_elem_per_dir
        ncp = patch.n_cp_per_dir
        strides = np.concatenate([[1], np.cumprod(ncp[:-1])]).astype(np.int64)
        eidx = np.empty((self.n_elems, self.dim), dtype=np.int64)
        rem = np.arange(self.n_elems)
        for d in range(self.dim):
            eidx[:, d] = rem % ne[d]
            rem = rem // ne[d]
        self.elem_index = eidx
        local = _block_indices([np.arange(p + 1) for p in self.degree], strides)
        base = sum(spans[d][2][eidx[:, d]] * strides[d] for d in range(self.dim))
        self.elem_dofs = (base[:, None] + local[None, :]).astype(np.int64)
        self.elem_vertex = np.concatenate(
            [np.stack([spans[d][0][eidx[:, d]] for d in range(self.dim)], 1),
             np.stack([spans[d][1][eidx[:, d]] for d in range(self.dim)], 1)], 1)
        self.boundaries = self._classify_boundaries()
        self._device = None
        self._patterns = {}
        self._cpts_t = None  # control points as a tensor carrying a gradient (replace)

    # -- reference-compatible properties -------------------------------------
    @property
    def n_bernstein(self) -> int:
        return int(np.prod([p + 1 for p in self.degree]))

    @property
    def n_local(self) -> int:
        return self.elem_dofs.shape[1]

    @property
    def n_dofs(self) -> int:
        return self.n_scalar_basis * self.vec

    @property
    def dim_phys(self) -> int:
        return self.cpts.shape[1]

    @property
    def extraction(self) -> np.ndarray:
        """Full per-element operators ``(n_e, n_l, n_l)``, later directions
        outermost in the Kronecker product (host, built on demand)."""
        out = np.empty((self.n_elems, self.n_local, self.n_local))
        for e in range(self.n_elems):
            C = self.extraction_ops[-1][self.elem_index[e, -1]]
            for d in range(self.dim - 2, -1, -1):
                C = np.kron(C, self.extraction_ops[d][self.elem_index[e, d]])
            out[e] = C
        return out

    def replace(self, *, cpts=None):
        """A copy on new control points (same knots, weights and connectivity),
        the role ``dataclasses.replace(V, cpts=cpts)`` plays in the reference's
        shape derivatives (``tests/test_galerkin.py:233-262``). A torch tensor
        that requires grad makes :func:`jaxiga_amd.solve` differentiable in
        the control points; the copy shares the CSR patterns (they depend on
        the connectivity only)."""
        import copy

        import torch

        new = copy.copy(self)
        if cpts is not None:
            if isinstance(cpts, torch.Tensor):
                arr = cpts.detach().to("cpu", torch.float64).numpy().copy()
                new._cpts_t = cpts
            else:
                arr = np.array(cpts, dtype=float)
                new._cpts_t = None
            if arr.shape != self.cpts.shape:
                raise ValueError(f"control points of shape {arr.shape}, expected {self.cpts.shape}")
            new.cpts = arr
            new.patch = dataclasses.replace(self.patch, ctrl_pts=arr)
        new._device = None
        return new

    def elem_dofs_vec(self) -> np.ndarray:
        return expand_dofs(self.elem_dofs, self.vec)

    def boundary(self, label: str) -> BoundarySet:
        if label not in self.boundaries:
            raise KeyError(f"unknown boundary label {label!r}; available: {sorted(self.boundaries)}")
        return self.boundaries[label]

    def summary(self) -> str:
        return (f"FunctionSpace(dim={self.dim}, vec={self.vec}, degree={self.degree}, "
                f"n_elems={self.n_elems}, n_scalar_basis={self.n_scalar_basis}, "
                f"n_dofs={self.n_dofs}, boundaries={sorted(self.boundaries)})")

    # -- setup ------------------------------------------------------------------
    def _classify_boundaries(self):
        ne = self.n_elem_per_dir
        estr = np.concatenate([[1], np.cumprod(ne[:-1])]).astype(np.int64)
        labels = self.patch.label_map
        collected = {}
        for d in range(self.dim):
            for end in (0, 1):
                side = side_name(d, end)
                label = labels.get(side, f"patch0/{side}")
                rng = [np.arange(n) for n in ne]
                rng[d] = np.array([0 if end == 0 else ne[d] - 1])
                elems = _block_indices(rng, estr)
                entry = collected.setdefault(label, ([], [], []))
                entry[0].append(elems)
                entry[1].append(np.full(len(elems), side_code(side), dtype=np.int64))
                entry[2].append(side_indices(self.patch, side))
        return {k: BoundarySet(np.concatenate(v[0]), np.concatenate(v[1]),
                               np.unique(np.concatenate(v[2])))
                for k, v in collected.items()}

    def neighbour_ranges(self):
        """Per direction, inclusive range of basis indices sharing a span with
        each basis index: ``[(lo (n_cp_d,), hi (n_cp_d,)), ...]``."""
        out = []
        for d in range(self.dim):
            p = self.degree[d]
            first = self._spans[d][2]
            n = self.patch.n_cp_per_dir[d]
            lo = np.full(n, np.iinfo(np.int64).max)
            hi = np.full(n, -1)
            for a in range(p + 1):
                np.minimum.at(lo, first + a, first)
                np.maximum.at(hi, first + a, first + p)
            out.append((lo, hi))
        return out

    def support_spans(self):
        """Per direction, the first and last span on which each 1-D basis
        function is supported: ``[(elo (n_cp_d,), ehi (n_cp_d,)), ...]``."""
        out = []
        for d in range(self.dim):
            p = self.degree[d]
            first = self._spans[d][2]
            n = self.patch.n_cp_per_dir[d]
            elo = np.full(n, np.iinfo(np.int64).max)
            ehi = np.full(n, -1)
            span = np.arange(len(first))
            for a in range(p + 1):
                np.minimum.at(elo, first + a, span)
                np.maximum.at(ehi, first + a, span)
            out.append((elo, ehi))
        return out

    # -- device tables ------------------------------------------------------------
    def device(self):
        """Lazily built :class:`DeviceSpace` (tables on the current GPU)."""
        if self._device is None:
            self._device = DeviceSpace(self)
        return self._device


def _regular_spans(space, sup) -> bool:
    """True when no interior knot repeats in any direction, so that span e
    starts at function e and function i is supported on spans
    max(0, i - p) .. min(n_elem - 1, i)."""
    p = space.degree[0]
    for d in range(space.dim):
        ne = space.n_elem_per_dir[d]
        n = space.patch.n_cp_per_dir[d]
        i = np.arange(n)
        elo, ehi = sup[d]
        if (n != ne + p or not np.array_equal(space._spans[d][2], np.arange(ne))
                or not np.array_equal(elo, np.maximum(0, i - p))
                or not np.array_equal(ehi, np.minimum(ne - 1, i))):
            return False
    return True


class DeviceSpace:
    """Patch tables in HBM plus the ``iga_patch_t`` that points at them."""

    def __init__(self, space: FunctionSpace):
        import torch

        from jaxiga_amd import _lib
        from jaxiga_amd.runtime import device

        dev = device()
        if len(set(space.degree)) != 1:
            raise NotImplementedError(f"GPU kernels need one degree in every direction, got {space.degree}")
        if space.dim_phys != space.dim:
            raise NotImplementedError("embedded surfaces are outside the supported hot path")
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.space_ref = weakref.ref(space)
        self.dev = dev
        self.ext = [torch.as_tensor(np.ascontiguousarray(op), **f64) for op in space.extraction_ops]
        self.first = [torch.as_tensor(s[2], **i32) for s in space._spans]
        self.lo = [torch.as_tensor(s[0], **f64) for s in space._spans]
        self.hi = [torch.as_tensor(s[1], **f64) for s in space._spans]
        self.cpts = torch.as_tensor(np.ascontiguousarray(space.cpts), **f64)
        self.wgts = torch.as_tensor(space.wgts, **f64)
        nb = space.neighbour_ranges()
        self.nb_lo = [torch.as_tensor(lo, **i32) for lo, _ in nb]
        self.nb_hi = [torch.as_tensor(hi, **i32) for _, hi in nb]
        sup = space.support_spans()
        self.elo = [torch.as_tensor(lo, **i32) for lo, _ in sup]
        self.ehi = [torch.as_tensor(hi, **i32) for _, hi in sup]
        self.elem_dofs = torch.as_tensor(space.elem_dofs, **i32)
        s = _lib.IgaPatch()
        s.dim = space.dim
        s.degree = space.degree[0]
        for d in range(3):
            if d < space.dim:
                s.n_elem[d] = space.n_elem_per_dir[d]
                s.n_cp[d] = space.patch.n_cp_per_dir[d]
                s.ext[d] = self.ext[d].data_ptr()
                s.first[d] = self.first[d].data_ptr()
                s.lo[d] = self.lo[d].data_ptr()
                s.hi[d] = self.hi[d].data_ptr()
                s.elo[d] = self.elo[d].data_ptr()
                s.ehi[d] = self.ehi[d].data_ptr()
            else:
                s.n_elem[d] = 1
                s.n_cp[d] = 1
        s.cpts = self.cpts.data_ptr()
        s.wgts = self.wgts.data_ptr()
        s.regular = int(_regular_spans(space, sup))
        self.struct = s


@dataclasses.dataclass(frozen=True, eq=False)
class DofMap:
    """Partition of the global dofs into free and prescribed
    (``function_space.py:554-603``)."""

    free: np.ndarray
    prescribed: np.ndarray
    values: np.ndarray
    n_dofs: int = 0

    @property
    def n_free(self) -> int:
        return len(self.free)

    @functools.cached_property
    def key(self) -> bytes:
        """Digest of the free set, from the (short) prescribed list: the free
        dofs are the sorted complement of the prescribed ones in
        ``range(n_dofs)``, so this identifies the free set without hashing it."""
        h = hashlib.sha1(np.int64(self.n_dofs).tobytes())
        h.update(np.ascontiguousarray(self.prescribed, dtype=np.int64).tobytes())
        return h.digest()

    def lift(self, u_free):
        import torch

        if isinstance(u_free, torch.Tensor):
            u = torch.zeros(self.n_dofs, dtype=u_free.dtype, device=u_free.device)
            u[torch.as_tensor(self.free, device=u.device)] = u_free
            if len(self.prescribed):
                u[torch.as_tensor(self.prescribed, device=u.device)] = torch.as_tensor(
                    self.values, dtype=u.dtype, device=u.device)
            return u
        u = np.zeros(self.n_dofs)
        u[self.free] = u_free
        u[self.prescribed] = self.values
        return u

    def restrict(self, v_full):
        return v_full[self.free]

    def zeros(self):
        return np.zeros(self.n_free)


def make_dofmap(n_dofs: int, prescribed, values) -> DofMap:
    prescribed = np.asarray(prescribed, dtype=np.int64)
    order = np.argsort(prescribed, kind="stable")
    prescribed = prescribed[order]
    values = np.asarray(values, float)[order] if len(prescribed) else np.zeros(0)
    free = np.setdiff1d(np.arange(n_dofs), prescribed)
    return DofMap(free=free, prescribed=prescribed, values=values, n_dofs=int(n_dofs))
