"""The element-list path: the padded-element contract and chunked assembly.

The reference's ``FunctionSpace`` carries, besides its tensor structure, the
element connectivity ``elem_dofs (n_e, n_local_max)`` and the per-element
Bezier extraction ``(n_e, n_local_max, n_bernstein)``; hierarchical spaces
pad short elements with the sentinel dof ``n_scalar_basis`` and zero
extraction rows (``space/function_space.py:23-38``). Its methods evaluate the
basis in chunks of elements so that memory is set by the chunk
(``ChunkedBasis``, ``space/evaluation.py:256-302``; ``methods/_common.py:125-160``),
and scatter-add element matrices through a slot map built from sorted unique
keys (``methods/_common.py:203-264``).

:class:`ElementSpace` is that contract on the GPU, and :func:`solve` the
Galerkin linear solve over it:

* the CSR pattern and the slot map come from ``iga_elem_pattern`` (keys,
  radix sort, unique, binary searches);
* each chunk of elements runs ``iga_elem_basis`` (basis data of the chunk
  only) and ``iga_elem_assemble`` (tangents and load, scatter-added with FP64
  atomics through the slot map, the lifting included);
* ``jaxiga_amd.solve(..., chunk=c)`` on a tensor space takes this path with
  the space's unpadded connectivity, which bounds the transient memory of
  assembly by the chunk.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd.runtime import device, lib, stream
from jaxiga_amd.space import make_dofmap


class ElementSpace:
    """A space given by element connectivity and per-element extraction, with
    the reference's padding: ``elem_dofs`` entries equal to
    ``n_scalar_basis`` and zero extraction rows mark absent functions."""

    cell_type = "tensor"

    def __init__(self, *, degree, cpts, wgts, elem_dofs, extraction, elem_vertex, vec=1, boundaries=None):
        self.degree = tuple(int(p) for p in degree)
        if len(set(self.degree)) != 1:
            raise NotImplementedError("the element-list kernels take one degree in every direction")
        self.dim = len(self.degree)
        self.vec = int(vec)
        self.cpts = np.asarray(cpts, float)
        self.wgts = np.asarray(wgts, float)
        self.elem_dofs = np.asarray(elem_dofs, np.int64)
        self.extraction = np.asarray(extraction, float)
        self.elem_vertex = np.asarray(elem_vertex, float)
        self.n_scalar_basis = self.cpts.shape[0]
        self.n_elems, self.n_local_max = self.elem_dofs.shape
        nb = (self.degree[0] + 1) ** self.dim
        if self.extraction.shape != (self.n_elems, self.n_local_max, nb):
            raise ValueError(f"extraction of shape {self.extraction.shape}, expected "
                             f"{(self.n_elems, self.n_local_max, nb)}")
        pad = self.elem_dofs == self.n_scalar_basis
        if np.any(self.elem_dofs > self.n_scalar_basis) or np.any(self.elem_dofs < 0):
            raise ValueError("elem_dofs entries must lie in [0, n_scalar_basis] (n_scalar_basis pads)")
        if np.any(self.extraction[pad] != 0.0):
            raise ValueError("extraction rows under the padding dof must be zero")
        self.boundaries = dict(boundaries or {})
        self._dev = None
        self._patterns = {}

    @property
    def n_dofs(self) -> int:
        return self.n_scalar_basis * self.vec

    @property
    def n_local(self) -> int:
        return self.n_local_max

    @property
    def dim_phys(self) -> int:
        return self.cpts.shape[1]

    def boundary(self, label):
        if label not in self.boundaries:
            raise KeyError(f"unknown boundary label {label!r}; available: {sorted(self.boundaries)}")
        return self.boundaries[label]

    @classmethod
    def from_space(cls, space, n_local_max=None, permute_seed=None):
        """The element-list view of a tensor :class:`FunctionSpace`: its full
        per-element extraction, padded to ``n_local_max`` functions with the
        sentinel (and zero rows), the local functions optionally permuted per
        element (``permute_seed``)."""
        ed = np.asarray(space.elem_dofs, np.int64)
        ext = space.extraction
        n_e, n_l = ed.shape
        m = n_l if n_local_max is None else int(n_local_max)
        if m < n_l:
            raise ValueError(f"n_local_max {m} is below the {n_l} functions of an element")
        dofs = np.full((n_e, m), space.n_scalar_basis, np.int64)
        C = np.zeros((n_e, m, ext.shape[2]))
        dofs[:, :n_l] = ed
        C[:, :n_l] = ext
        if permute_seed is not None:
            rng = np.random.default_rng(permute_seed)
            perm = np.argsort(rng.random((n_e, m)), axis=1)
            dofs = np.take_along_axis(dofs, perm, 1)
            C = np.take_along_axis(C, perm[:, :, None], 1)
        return cls(degree=space.degree, cpts=space.cpts, wgts=space.wgts, elem_dofs=dofs, extraction=C,
                   elem_vertex=space.elem_vertex, vec=space.vec, boundaries=space.boundaries)

    def device_arrays(self):
        import torch

        if self._dev is None:
            dev = device()
            f64 = dict(dtype=torch.float64, device=dev)
            self._dev = {
                "elem_dofs": torch.as_tensor(self.elem_dofs.astype(np.int32), device=dev),
                "ext": torch.as_tensor(np.ascontiguousarray(self.extraction), **f64),
                "cpts": torch.as_tensor(np.ascontiguousarray(self.cpts), **f64),
                "wgts": torch.as_tensor(self.wgts, **f64),
                "boxes": torch.as_tensor(np.ascontiguousarray(self.elem_vertex), **f64),
            }
        return self._dev


def element_pattern(espace, dofmap, f2f_d):
    """The CSR pattern of the free-free block with the slot map ``scatter``
    ``(n_e, n_l*vec, n_l*vec)`` (nnz for padding and prescribed dofs)."""
    import torch

    key = (espace.vec, dofmap.key)
    pat = espace._patterns.get(key)
    if pat is not None:
        return pat
    L = lib()
    d = espace.device_arrays()
    n_free = len(dofmap.free)
    n_e, m, vec = espace.n_elems, espace.n_local_max, espace.vec
    nbytes = int(L.iga_elem_pattern_workspace(n_e, m, vec))
    if nbytes < 0:
        raise ValueError("the element-pair count exceeds 2^31 - 1; use a coarser mesh")
    dev = f2f_d.device
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    nnz = _lib.c_i64(0)
    _lib.check(L.iga_elem_pattern(n_e, m, vec, _lib.ptr(d["elem_dofs"]), espace.n_scalar_basis, _lib.ptr(f2f_d),
                                  n_free, _lib.ptr(work), nbytes, None, None, None, None,
                                  _lib.ctypes.byref(nnz), stream()), "iga_elem_pattern")
    i32 = dict(dtype=torch.int32, device=dev)
    indptr = torch.empty(n_free + 1, **i32)
    cols = torch.empty(max(nnz.value, 1), **i32)[:nnz.value]
    rows = torch.empty(max(nnz.value, 1), **i32)[:nnz.value]
    scatter = torch.empty((n_e, m * vec, m * vec), **i32)
    _lib.check(L.iga_elem_pattern(n_e, m, vec, _lib.ptr(d["elem_dofs"]), espace.n_scalar_basis, _lib.ptr(f2f_d),
                                  n_free, _lib.ptr(work), nbytes, _lib.ptr(indptr), _lib.ptr(cols),
                                  _lib.ptr(rows), _lib.ptr(scatter), _lib.ctypes.byref(nnz), stream()),
               "iga_elem_pattern")
    pat = LIN.CSRPattern(indptr, cols, rows, n_free)
    pat.scatter_d = scatter
    espace._patterns[key] = pat
    return pat


@dataclasses.dataclass
class ElementContext:
    problem: object
    dofmap: object
    n_q1: int
    free_d: object
    f2f_d: object
    gfull_d: object
    rule_pts: object
    rule_wts: object
    pattern: object

    @property
    def space(self):
        return self.problem.space

    @property
    def n_free(self):
        return len(self.dofmap.free)

    @property
    def assembly(self):
        return self.pattern

    def rule(self):
        r = _lib.IgaRule()
        r.n_q1 = self.n_q1
        r.pts = self.rule_pts.data_ptr()
        r.wts = self.rule_wts.data_ptr()
        return r


def build_context(problem, espace, params=None):
    import torch

    dev = device()
    prescribed, values = problem.resolve_dirichlet(params)
    dofmap = make_dofmap(espace.n_dofs, prescribed, values)
    f2f = np.full(espace.n_dofs, -1, dtype=np.int32)
    f2f[dofmap.free] = np.arange(len(dofmap.free), dtype=np.int32)
    gfull = np.zeros(espace.n_dofs)
    gfull[dofmap.prescribed] = dofmap.values
    n1 = espace.degree[0] + 1
    pts, wts = np.polynomial.legendre.leggauss(n1)
    f64 = dict(dtype=torch.float64, device=dev)
    f2f_d = torch.as_tensor(f2f, device=dev)
    return ElementContext(problem=problem, dofmap=dofmap, n_q1=n1,
                          free_d=torch.as_tensor(dofmap.free.astype(np.int32), device=dev), f2f_d=f2f_d,
                          gfull_d=torch.as_tensor(gfull, **f64), rule_pts=torch.as_tensor(pts, **f64),
                          rule_wts=torch.as_tensor(wts, **f64),
                          pattern=element_pattern(espace, dofmap, f2f_d))


def _per_point(value, params, x, lo, hi, vec, kind):
    """A coefficient or source for a chunk: None, a constant ``(vec,)``, or
    values at the chunk's points ``(n_c, n_q[, vec])``."""
    import torch

    from jaxiga_amd.forms import PointField
    from jaxiga_amd.methods import call_points

    if value is None:
        return None
    if isinstance(value, PointField):
        vals = torch.as_tensor(value.values, dtype=torch.float64, device=x.device)[lo:hi]
        return vals.contiguous()
    if callable(value):
        out = call_points(value, x, params=params, vec=vec)
        return out[..., 0].contiguous() if kind == "coef" else out.contiguous()
    if isinstance(value, torch.Tensor):
        value = value.detach().cpu().numpy()
    return np.broadcast_to(np.asarray(value, float), (vec,))


def assemble(ctx: ElementContext, espace: ElementSpace, params=None, chunk=None, values=True, load=True):
    """CSR values and the free load, assembled chunk by chunk (``chunk``
    elements at a time; all at once when None)."""
    import torch

    L = lib()
    problem = ctx.problem
    d = espace.device_arrays()
    dev = ctx.f2f_d.device
    f64 = dict(dtype=torch.float64, device=dev)
    pat = ctx.pattern
    vals = torch.zeros(pat.nnz, **f64) if values else None
    rhs = torch.zeros(ctx.n_free, **f64) if load else None
    n_e, m, dim, vec = espace.n_elems, espace.n_local_max, espace.dim, espace.vec
    n_q = ctx.n_q1 ** dim
    step = n_e if chunk is None else max(1, int(chunk))
    args = _lib.IgaAsmArgs()
    args.vec = vec
    args.precision = _lib.IGA_FP64
    if problem.kernel == "poisson":
        args.problem = _lib.IGA_POISSON
        a0 = params.get("a0") if isinstance(params, dict) else None
    elif problem.kernel == "elasticity":
        args.problem = _lib.IGA_ELASTICITY
        C = problem.C(params)
        C = C.detach().cpu().numpy() if isinstance(C, torch.Tensor) else np.asarray(C, float)
        for i, c in enumerate(C.ravel()):
            args.cmat[i] = float(c)
        a0 = None
    else:
        raise NotImplementedError(f"{type(problem).__name__} has no element-list kernel")
    args.full_to_free = ctx.f2f_d.data_ptr()
    args.gfull = ctx.gfull_d.data_ptr()
    args.nnz = pat.nnz
    args.vals = vals.data_ptr() if values and pat.nnz else None
    args.rhs = rhs.data_ptr() if load and ctx.n_free else None
    need_x = callable(a0) or callable(problem.source)
    for lo in range(0, n_e, step):
        hi = min(n_e, lo + step)
        n_c = hi - lo
        R = torch.empty((n_c, n_q, m), **f64)
        dR = torch.empty((n_c, n_q, dim, m), **f64)
        x = torch.empty((n_c, n_q, dim), **f64) if need_x else None
        w = torch.empty((n_c, n_q), **f64)
        _lib.check(L.iga_elem_basis(dim, espace.degree[0], m, lo, n_c, ctx.rule(), _lib.ptr(d["ext"]),
                                    _lib.ptr(d["elem_dofs"]), espace.n_scalar_basis, _lib.ptr(d["cpts"]),
                                    _lib.ptr(d["wgts"]), _lib.ptr(d["boxes"]), _lib.ptr(R), _lib.ptr(dR),
                                    _lib.ptr(x), _lib.ptr(w), stream()), "iga_elem_basis")
        keep = []
        coef = _per_point(a0, params, x, lo, hi, 1, "coef")
        if coef is None or not isinstance(coef, torch.Tensor):
            args.coef_kind = _lib.IGA_COEF_CONST
            args.coef_const = 1.0 if coef is None else float(np.asarray(coef).ravel()[0])
        else:
            args.coef_kind, args.coef_points = _lib.IGA_COEF_POINTS, coef.data_ptr()
            keep.append(coef)
        src = _per_point(problem.source, params, x, lo, hi, vec, "src") if load else None
        if src is None:
            args.src_kind = _lib.IGA_SRC_NONE
        elif isinstance(src, torch.Tensor):
            src = src.expand(n_c, n_q, vec).contiguous()
            args.src_kind, args.src_points = _lib.IGA_SRC_POINTS, src.data_ptr()
            keep.append(src)
        else:
            args.src_kind = _lib.IGA_SRC_CONST
            for c in range(vec):
                args.src_const[c] = float(src[c])
        _lib.check(L.iga_elem_assemble(dim, n_q, m, lo, n_c, _lib.ptr(R), _lib.ptr(dR), _lib.ptr(w),
                                       _lib.ptr(d["elem_dofs"]), espace.n_scalar_basis, args,
                                       _lib.ptr(pat.scatter_d), stream()), "iga_elem_assemble")
        del keep
    return vals, rhs


def _carries_gradient(params, space):
    import torch

    from jaxiga_amd.forms import PointField

    vals = list(params.values()) if isinstance(params, dict) else []
    vals = [v.values if isinstance(v, PointField) else v for v in vals]
    vals.append(getattr(space, "_cpts_t", None))
    return any(isinstance(v, torch.Tensor) and v.requires_grad for v in vals)


def solve(problem, params=None, linear=None, chunk=None):
    """Galerkin linear solve over the element-list path; ``(u, ctx, stats)``
    as :func:`jaxiga_amd.methods.solve_galerkin` returns them."""
    import torch

    from jaxiga_amd import methods as M

    space = problem.space
    if torch.is_grad_enabled() and _carries_gradient(params, space):
        raise NotImplementedError("the element-list path is not differentiated; solve without chunk= "
                                  "on the tensor space for gradients")
    tensor = not isinstance(space, ElementSpace)
    espace = ElementSpace.from_space(space) if tensor else space
    if problem.neumann and not tensor:
        raise NotImplementedError("Neumann loads on an element-list space are not provided")
    if chunk is not None and (isinstance(chunk, bool) or int(chunk) <= 0):
        raise ValueError(f"chunk must be a positive number of elements, got {chunk!r}")
    t0 = time.perf_counter()
    ctx = build_context(problem, espace, params)
    if linear is None:
        linear = M.default_linear(problem, ctx.n_free)
    if linear.matrix_free or linear.preconditioner == "multigrid":
        raise NotImplementedError("the element-list path solves the assembled operator with "
                                  "Jacobi/none preconditioning")
    vals, b = assemble(ctx, espace, params, chunk)
    if problem.neumann:
        M.neumann_load(M.build_context(problem, params), params, b)
    u_free, info = LIN.sparse_solve(linear, ctx.pattern, vals, b)
    rn, bn = LIN.residual_norms(ctx.pattern, vals, u_free, b)
    torch.cuda.synchronize()
    stats = {"method": "galerkin", "linear": linear.method, "matrix_free": False, "n_free": ctx.n_free,
             "nnz": ctx.pattern.nnz, **info, "chunk": chunk, "wall_time": time.perf_counter() - t0}
    stats.update(LIN.linear_diagnostics(rn, bn, u_free, linear.tol))
    LIN.warn_if_failed("linear Galerkin solve", stats["converged"], stats["residual"])
    return ctx.dofmap.lift(u_free), ctx, stats
