"""Oracle assembly: CSR pattern, scatter-add, load vector with lifting.

Restates
* ``src/jaxiga/methods/_common.py`` -- _cached_csr_pattern :203-264 (unique
  free-free keys ``row*n + col``, searchsorted scatter map with an ``nnz``
  sentinel, int32 indices, ``indptr`` by bincount);
* ``src/jaxiga/methods/galerkin.py`` -- _assemble_csr_values :122-142
  (``values[scatter] += K_e``, sentinel dropped), rhs :240-245 (``b = -r(0)``,
  i.e. source + Neumann work - K_fp g);
* ``src/jaxiga/solvers/linear.py`` -- CSRPattern :38-68.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from oracle.space import expand_dofs


@dataclasses.dataclass
class CSRPattern:
    indptr: np.ndarray
    column_indices: np.ndarray
    row_indices: np.ndarray
    scatter: np.ndarray  # (n_e, n_l*vec, n_l*vec) int32, sentinel nnz
    shape: tuple

    @property
    def nnz(self):
        return len(self.column_indices)


def csr_pattern(elem_dofs, n_dofs, vec, free, chunk=4_000_000):
    free = np.asarray(free, dtype=np.int64)
    n = len(free)
    to_free = np.full(n_dofs + vec, -1, dtype=np.int64)
    to_free[free] = np.arange(n)
    loc = to_free[expand_dofs(np.asarray(elem_dofs), vec)]
    n_e, n_l = loc.shape
    step = max(1, chunk // max(n_l * n_l, 1))

    def keys_of(lo):
        blk = loc[lo : lo + step]
        return blk[:, :, None] * n + blk[:, None, :], (blk[:, :, None] >= 0) & (blk[:, None, :] >= 0)

    parts = []
    for lo in range(0, n_e, step):
        k, v = keys_of(lo)
        parts.append(np.unique(k[v]))
    keys = np.unique(np.concatenate(parts)) if parts else np.zeros(0, np.int64)
    if len(keys) >= np.iinfo(np.int32).max:
        raise OverflowError("too many nonzeros for int32")
    scatter = np.full((n_e, n_l, n_l), len(keys), dtype=np.int32)
    for lo in range(0, n_e, step):
        k, v = keys_of(lo)
        blk = scatter[lo : lo + step]
        blk[v] = np.searchsorted(keys, k[v]).astype(np.int32)
    rows = (keys // max(n, 1)).astype(np.int32)
    cols = (keys % max(n, 1)).astype(np.int32)
    counts = np.bincount(rows, minlength=n)
    indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return CSRPattern(indptr, cols, rows, scatter, (n, n))


def assemble_csr(pattern, Ke):
    vals = np.bincount(pattern.scatter.ravel(), weights=np.asarray(Ke).ravel(), minlength=pattern.nnz + 1)
    return vals[: pattern.nnz]


def global_load(n_dofs, gdofs, Fe):
    """Scatter element loads ``(n_e, n_loc)`` into a full dof vector."""
    return np.bincount(np.asarray(gdofs).ravel(), weights=np.asarray(Fe).ravel(), minlength=n_dofs)[:n_dofs]


def lifting(n_dofs, gdofs, Ke, dofmap):
    """``(K u_g)`` over all dofs, with ``u_g`` zero on free dofs and the
    prescribed values elsewhere."""
    ug = np.zeros(n_dofs)
    ug[dofmap.prescribed] = dofmap.values
    ue = ug[gdofs]  # (n_e, n_loc)
    return global_load(n_dofs, gdofs, np.einsum("eij,ej->ei", Ke, ue))


def csr_matvec(pattern, data, v):
    out = np.zeros(pattern.shape[0])
    np.add.at(out, pattern.row_indices, data * v[pattern.column_indices])
    return out


def to_scipy(pattern, data):
    import scipy.sparse as sp

    return sp.csr_matrix((data, pattern.column_indices, pattern.indptr), shape=pattern.shape)
