"""Linear-solver options, the device CSR pattern, and GPU Krylov solves.

Mirrors ``solvers/linear.py`` of the reference: ``CSRPattern`` :38-68,
``LinearOptions`` :76-127 (same fields, defaults and validation),
``_iterative_csr`` :324-351 (Jacobi-preconditioned CG from zero, stop at
``||r|| <= tol ||b||``), ``sparse_solve`` :392-396, ``matrix_free_solve``
:414-442; and ``solvers/diagnostics.py`` :8-34.

The CSR pattern of the free-free block is generated on the GPU from the
tensor structure (row ``r`` lists the free dofs of the tensor neighbourhood of
its scalar basis function, in increasing order), which is the same sorted,
unique pattern the reference builds with ``np.unique`` over element keys.
"""

from __future__ import annotations

import dataclasses
import warnings

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd.runtime import device, lib, stream

SOLVE_COUNTER = {"forward": 0, "adjoint": 0}


def reset_solve_counter():
    SOLVE_COUNTER["forward"] = 0
    SOLVE_COUNTER["adjoint"] = 0


@dataclasses.dataclass(frozen=True)
class LinearOptions:
    """Linear-solver configuration (reference field names and checks).

    On this backend ``"cg"`` runs the device-resident Jacobi-PCG;
    ``"dense"`` runs a dense LU on the GPU for systems up to ``DIRECT_LIMIT``
    free dofs; ``"scipy"`` is the reference's host SuperLU (any size). ``preconditioner="multigrid"``
    (an addition; the reference allows ``"none"`` and ``"jacobi"``) runs
    flexible CG with one multigrid V-cycle per iteration through
    :func:`jaxiga_amd.solve` -- damped-Jacobi smoothers, or the smoothers of
    the generator registered with :func:`jaxiga_amd.learned.use_generator` --
    on scalar 2-D problems whose prescribed dofs are whole sides.
    """

    method: str = "scipy"
    matrix_free: bool = False
    preconditioner: str = "jacobi"
    tol: float = 1e-10
    maxiter: int | None = None
    superlu_ordering: str = "auto"

    def __post_init__(self):
        if not np.isfinite(self.tol) or self.tol <= 0:
            raise ValueError("tol must be finite and positive")
        if self.maxiter is not None and (
            isinstance(self.maxiter, (bool, np.bool_))
            or not isinstance(self.maxiter, (int, np.integer)) or self.maxiter <= 0
        ):
            raise ValueError("maxiter must be a positive integer or None")
        allowed = {"cg", "bicgstab", "gmres", "dense", "scipy"}
        if self.method not in allowed:
            raise ValueError(f"unknown linear method {self.method!r}; expected one of {allowed}")
        if self.preconditioner not in {"none", "jacobi", "multigrid"}:
            raise ValueError(f"unknown preconditioner {self.preconditioner!r}; expected 'none', 'jacobi' "
                             "or 'multigrid'")
        if self.preconditioner == "multigrid" and (self.method != "cg" or self.matrix_free):
            raise ValueError("preconditioner='multigrid' runs flexible CG on the assembled operator: "
                             "use method='cg' with matrix_free=False")
        orderings = {"auto", "NATURAL", "MMD_ATA", "MMD_AT_PLUS_A", "COLAMD"}
        if self.superlu_ordering not in orderings:
            raise ValueError(f"unknown SuperLU ordering {self.superlu_ordering!r}; expected one of {orderings}")
        if self.matrix_free and self.method in {"dense", "scipy"}:
            raise ValueError(f"matrix_free=True needs an iterative method, not {self.method!r}; "
                             "there is no matrix for a direct solver to factor")


DIRECT_LIMIT = 20000


class CSRPattern:
    """Static CSR structure of the free-free block, resident on the GPU."""

    def __init__(self, indptr, cols, rows, n):
        self.indptr_d = indptr
        self.cols_d = cols
        self.rows_d = rows
        self.shape = (n, n)
        self.symmetric = True
        self._host = None

    @property
    def nnz(self) -> int:
        return int(self.cols_d.shape[0])

    def _h(self):
        if self._host is None:
            self._host = (self.indptr_d.cpu().numpy(), self.cols_d.cpu().numpy(),
                          self.rows_d.cpu().numpy())
        return self._host

    @property
    def indptr(self):
        return self._h()[0]

    @property
    def column_indices(self):
        return self._h()[1]

    @property
    def row_indices(self):
        return self._h()[2]

    @property
    def coo_indices(self):
        return np.stack([self.row_indices, self.column_indices], axis=1)


def build_pattern(space, free_dev, f2f_dev, n_free) -> CSRPattern:
    """Generate the CSR pattern on the GPU (count, scan, fill)."""
    import torch

    ds = space.device()
    L = lib()
    st = stream()
    dev = ds.dev
    counts = torch.empty(max(n_free, 1), dtype=torch.int32, device=dev)
    ncp = (_lib.c_int * 3)(*space.patch.n_cp_per_dir, *([1] * (3 - space.dim)))
    lo = (_lib.c_void_p * 3)(*[t.data_ptr() for t in ds.nb_lo], *([None] * (3 - space.dim)))
    hi = (_lib.c_void_p * 3)(*[t.data_ptr() for t in ds.nb_hi], *([None] * (3 - space.dim)))
    _lib.check(L.iga_csr_count(space.dim, ncp, lo, hi, space.vec, n_free, _lib.ptr(free_dev),
                               _lib.ptr(f2f_dev), _lib.ptr(counts), st), "iga_csr_count")
    indptr = torch.empty(n_free + 1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(L.iga_scan_workspace(n_free)), dtype=torch.int32, device=dev)
    _lib.check(L.iga_scan_exclusive(n_free, _lib.ptr(counts), _lib.ptr(indptr), _lib.ptr(ws), st),
               "iga_scan_exclusive")
    nnz = int(indptr[-1].item())
    cols = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    rows = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    _lib.check(L.iga_csr_fill(space.dim, ncp, lo, hi, space.vec, n_free, _lib.ptr(free_dev),
                              _lib.ptr(f2f_dev), _lib.ptr(indptr), _lib.ptr(cols), _lib.ptr(rows), st),
               "iga_csr_fill")
    return CSRPattern(indptr, cols, rows, n_free)


# ---------------------------------------------------------------------------
# operators and solves
# ---------------------------------------------------------------------------


def spmv(pattern: CSRPattern, data, x):
    import torch

    y = torch.empty_like(x)
    n = pattern.shape[0]
    _lib.check(lib().iga_spmv(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d), _lib.ptr(data),
                              _lib.ptr(x), _lib.ptr(y), stream()), "iga_spmv")
    return y


def inverse_diagonal(pattern: CSRPattern, data):
    import torch

    n = pattern.shape[0]
    inv = torch.empty(n, dtype=torch.float64, device=data.device)
    _lib.check(lib().iga_csr_diag_inv(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d),
                                      _lib.ptr(data), _lib.ptr(inv), stream()), "iga_csr_diag_inv")
    return inv


@dataclasses.dataclass
class CGResult:
    x: object
    iterations: int
    rr: float
    bb: float


def cg_csr(pattern: CSRPattern, data, b, tol=1e-10, maxiter=None, dinv=None, workspace=None,
           single_reduction=False):
    """Jacobi-PCG on the device; returns :class:`CGResult`. ``single_reduction``
    runs the Chronopoulos-Gear form (two kernels per iteration, the same
    iterates in exact arithmetic) instead of the textbook recurrence."""
    import torch

    L = lib()
    n = pattern.shape[0]
    if dinv is None:
        dinv = inverse_diagonal(pattern, data)
    if workspace is None:
        workspace = torch.empty(int(L.iga_cg_workspace(n)), dtype=torch.float64, device=b.device)
    x = torch.empty(n, dtype=torch.float64, device=b.device)
    it = _lib.c_int(0)
    rr = _lib.c_double(0.0)
    bb = _lib.c_double(0.0)
    maxiter = 10 * n if maxiter is None else int(maxiter)
    prec = _lib.IGA_FP32 if data.dtype == torch.float32 else _lib.IGA_FP64
    fn = L.iga_cg1_solve if single_reduction else L.iga_cg_solve
    _lib.check(fn(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d), _lib.ptr(data),
                  prec, _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x), float(tol), maxiter,
                  _lib.ptr(workspace), _lib.ctypes.byref(it), _lib.ctypes.byref(rr),
                  _lib.ctypes.byref(bb), stream()), "iga_cg_solve")
    return CGResult(x, it.value, rr.value, bb.value)


NOALLOC_MAX_N = 400000  # IGA_CG_NOALLOC_MAX_N of the C ABI


def cg_csr_noalloc(pattern: CSRPattern, data, b, x, dinv, tol=1e-10, maxiter=None):
    """Jacobi-PCG as :func:`cg_csr`, through ``iga_cg_solve_noalloc``: the
    library allocates nothing (its workspace is a fixed per-GPU device
    array), the caller passes ``x`` (written) and ``dinv``. Systems above
    ``NOALLOC_MAX_N`` rows are rejected. Calls that overlap on one GPU share
    the array and race."""
    L = lib()
    n = pattern.shape[0]
    if n > NOALLOC_MAX_N:
        raise ValueError(f"iga_cg_solve_noalloc takes at most {NOALLOC_MAX_N} rows, got {n}")
    it, rr, bb = _lib.c_int(0), _lib.c_double(0.0), _lib.c_double(0.0)
    maxiter = 10 * n if maxiter is None else int(maxiter)
    import torch

    prec = _lib.IGA_FP32 if data.dtype == torch.float32 else _lib.IGA_FP64
    _lib.check(L.iga_cg_solve_noalloc(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d),
                                      _lib.ptr(data), prec, _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x),
                                      float(tol), maxiter, _lib.ctypes.byref(it), _lib.ctypes.byref(rr),
                                      _lib.ctypes.byref(bb), stream()), "iga_cg_solve_noalloc")
    return CGResult(x, it.value, rr.value, bb.value)


def residual_norms(pattern: CSRPattern, data, u, b):
    """``(||A u - b||, ||b||)`` in FP64 on the device."""
    import torch

    L = lib()
    n = pattern.shape[0]
    out = torch.zeros(2, dtype=torch.float64, device=b.device)
    work = torch.empty(int(L.iga_cg_workspace(n)), dtype=torch.float64, device=b.device)
    _lib.check(L.iga_residual_norms(n, _lib.ptr(pattern.indptr_d), _lib.ptr(pattern.cols_d),
                                    _lib.ptr(data), _lib.ptr(u), _lib.ptr(b), _lib.ptr(out),
                                    _lib.ptr(work), stream()), "iga_residual_norms")
    o = out.cpu().numpy()
    return float(np.sqrt(o[0])), float(np.sqrt(o[1]))


def dense_from_csr(pattern: CSRPattern, data):
    import torch

    n = pattern.shape[0]
    A = torch.zeros((n, n), dtype=data.dtype, device=data.device)
    A[pattern.rows_d.long(), pattern.cols_d.long()] = data
    return A


def _solve_impl(opts: LinearOptions, pattern: CSRPattern, data, b):
    """``A u = b`` on the GPU with the method of ``opts``; ``(u, info)``."""
    import torch

    if opts.preconditioner == "multigrid":
        raise ValueError("preconditioner='multigrid' needs the problem's multigrid hierarchy: "
                         "solve through jaxiga_amd.solve")
    if opts.method == "cg":
        if opts.preconditioner == "jacobi":
            res = cg_csr(pattern, data, b, opts.tol, opts.maxiter)
        else:
            ones = torch.ones(pattern.shape[0], dtype=torch.float64, device=b.device)
            res = cg_csr(pattern, data, b, opts.tol, opts.maxiter, dinv=ones)
        return res.x, {"iterations": res.iterations}
    if opts.method == "dense":
        if pattern.shape[0] > DIRECT_LIMIT:
            raise ValueError(f"dense solves are limited to {DIRECT_LIMIT} free dofs on this "
                             "backend; use LinearOptions(method='cg') or method='scipy'")
        return torch.linalg.solve(dense_from_csr(pattern, data), b), {"iterations": None}
    if opts.method == "scipy":
        return _superlu_solve(opts, pattern, data, b), {"iterations": None}
    raise NotImplementedError(f"linear method {opts.method!r} is not provided on this backend; use 'cg'")


def _superlu_solve(opts: LinearOptions, pattern: CSRPattern, data, b):
    """``method="scipy"``: SuperLU on the host, as the reference runs it
    (``_scipy_csr_solve``, ``src/jaxiga/solvers/linear.py:221-250``, a
    ``pure_callback`` to SciPy): the device CSR values and load are copied to
    the host, factored with the symmetric ordering ``MMD_AT_PLUS_A`` in
    SuperLU's symmetric mode for ``superlu_ordering="auto"`` on the symmetric
    pattern (``spsolve`` with the named ordering otherwise), and the solution
    is copied back. An explicitly requested host direct solver, not a
    fallback: the device solvers are ``"cg"`` and ``"dense"``."""
    import scipy.sparse
    import scipy.sparse.linalg
    import torch

    n = pattern.shape[0]
    A = scipy.sparse.csr_matrix((data.detach().to(torch.float64).cpu().numpy(),
                                 pattern.column_indices, pattern.indptr), shape=pattern.shape)
    rhs = b.detach().cpu().numpy()
    ordering = opts.superlu_ordering
    if ordering == "auto":
        ordering = "MMD_AT_PLUS_A" if pattern.symmetric else "COLAMD"
    if n == 0:
        u = np.zeros(0)
    elif pattern.symmetric and ordering == "MMD_AT_PLUS_A":
        factor = scipy.sparse.linalg.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A",
                                          options={"SymmetricMode": True})
        u = factor.solve(rhs)
    else:
        u = scipy.sparse.linalg.spsolve(A, rhs, permc_spec=ordering)
    return torch.as_tensor(np.asarray(u, np.float64), device=b.device)


def csr_outer(pattern: CSRPattern, a, b, alpha=1.0):
    """``alpha * a[rows] * b[cols]`` over the pattern's nonzeros (FP64)."""
    import torch

    out = torch.empty(pattern.nnz, dtype=torch.float64, device=a.device)
    _lib.check(lib().iga_csr_outer(pattern.nnz, _lib.ptr(pattern.rows_d), _lib.ptr(pattern.cols_d),
                                   _lib.ptr(a), _lib.ptr(b), float(alpha), _lib.ptr(out), stream()),
               "iga_csr_outer")
    return out


def adjoint_solve(opts: LinearOptions, pattern: CSRPattern, data, u_bar):
    """The one adjoint solve of a reverse pass, ``A lam = u_bar`` on the same
    (symmetric) CSR; counted in ``SOLVE_COUNTER['adjoint']``."""
    if not pattern.symmetric:
        raise NotImplementedError("the CSR adjoint requires a symmetric pattern")
    SOLVE_COUNTER["adjoint"] += 1
    return _solve_impl(opts, pattern, data, u_bar.contiguous())[0]


def _sparse_solve_fn():
    import torch

    class SparseSolve(torch.autograd.Function):
        """``u = A(data)^-1 b`` with the implicit reverse rule of
        ``_csr_sparse_solve`` (``src/jaxiga/solvers/linear.py:362-389``):
        one adjoint solve ``A lam = u_bar`` on the same symmetric CSR, then
        ``data_bar = -lam[rows] u[cols]`` and ``b_bar = lam``."""

        @staticmethod
        def forward(fc, opts, pattern, info, data, b):
            u, inf = _solve_impl(opts, pattern, data.detach(), b.detach())
            info.update(inf)
            fc.opts, fc.pattern = opts, pattern
            fc.save_for_backward(data.detach(), u)
            return u

        @staticmethod
        def backward(fc, u_bar):
            data, u = fc.saved_tensors
            lam = adjoint_solve(fc.opts, fc.pattern, data, u_bar)
            data_bar = None
            if fc.needs_input_grad[3]:
                data_bar = csr_outer(fc.pattern, lam, u, -1.0).to(data.dtype)
            return None, None, None, data_bar, lam if fc.needs_input_grad[4] else None

    return SparseSolve


_SPARSE_SOLVE = []


def sparse_solve(opts: LinearOptions, pattern: CSRPattern, data, b):
    """Solve ``A u = b`` on the GPU; returns ``(u, info)``. Differentiable in
    ``data`` and ``b`` (one adjoint solve in the reverse pass, never through
    the iterations), as ``sparse_solve`` of ``src/jaxiga/solvers/linear.py:392-396``."""
    import torch

    SOLVE_COUNTER["forward"] += 1
    if torch.is_grad_enabled() and (data.requires_grad or b.requires_grad):
        if not _SPARSE_SOLVE:
            _SPARSE_SOLVE.append(_sparse_solve_fn())
        info = {}
        u = _SPARSE_SOLVE[0].apply(opts, pattern, info, data, b)
        return u, info
    return _solve_impl(opts, pattern, data, b)


def linear_diagnostics(residual_norm, b_norm, u, tol):
    import torch

    rel = residual_norm / (b_norm if b_norm > 0 else 1.0)
    finite = bool(np.isfinite(residual_norm) and bool(torch.isfinite(u).all()))
    return {"residual": residual_norm, "relative_residual": rel,
            "converged": finite and residual_norm <= tol * b_norm, "finite": finite}


def warn_if_failed(name, converged, residual):
    if not converged:
        warnings.warn(f"{name} did not converge (residual {residual:.3e}); "
                      "do not use this result for implicit sensitivities", RuntimeWarning, stacklevel=3)
