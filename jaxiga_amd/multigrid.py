"""Geometric multigrid for scalar 2-D tensor-patch operators.

The reference has no multigrid (its preconditioners are ``"none"`` and
``"jacobi"``, ``src/jaxiga/solvers/linear.py:113-116``); this is the backbone
of the learned preconditioner.

Levels come from dyadic coarsening of the knot vectors: every other distinct
interior knot is removed, so each coarse spline space is nested in the finer
one and the transfer ``P`` (coarse -> fine coefficients, the knot-insertion
matrix) is exact. The coarse operators are the Galerkin products
``P^T K P``, formed on the device by two masked sparse products over
patterns computed once per space on the host (like the reference's cached CSR
pattern, ``src/jaxiga/methods/_common.py:181-264``). The prescribed dofs must
be whole sides of the patch, so the free dofs of every level form a
rectangle of the control grid -- the ``h x w`` grid the stencil smoothers
run on (row-major, direction 0 fastest). The coarsest level is solved with a
dense inverse.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import scipy.sparse as sp

from jaxiga_amd import _lib
from jaxiga_amd.geometry import change_of_basis
from jaxiga_amd.runtime import lib, side_stream, stream


def coarsen_knots(knots, p):
    """Remove every other distinct interior knot (with its multiplicity);
    ``None`` when the number of spans is odd or below 4."""
    t = np.asarray(knots, float)
    u, mult = np.unique(t, return_counts=True)
    m = len(u) - 1
    if m % 2 or m < 4:
        return None
    keep = np.ones(len(u), bool)
    keep[1:-1:2] = False
    return np.repeat(u[keep], mult[keep])


def free_rectangle(n_cp, prescribed):
    """``(lo, hi)`` per direction of the free control points when the
    prescribed set is a union of whole sides; raises otherwise."""
    n0, n1 = n_cp
    pres = np.zeros(n0 * n1, bool)
    pres[np.asarray(prescribed, np.int64)] = True
    grid = pres.reshape(n1, n0)
    fixed = {(0, 0): grid[:, 0].all(), (0, 1): grid[:, -1].all(),
             (1, 0): grid[0, :].all(), (1, 1): grid[-1, :].all()}
    lo = [1 if fixed[(d, 0)] else 0 for d in range(2)]
    hi = [n_cp[d] - 2 if fixed[(d, 1)] else n_cp[d] - 1 for d in range(2)]
    rect = np.zeros_like(grid)
    rect[lo[1]: hi[1] + 1, lo[0]: hi[0] + 1] = True
    if not np.array_equal(~grid, rect):
        raise NotImplementedError("multigrid needs the prescribed dofs to be whole sides of the patch")
    return lo, hi, fixed


def _csr_dev(A, dev):
    import torch

    A = sp.csr_matrix(A)
    A.sort_indices()
    i32 = dict(dtype=torch.int32, device=dev)
    indptr = torch.as_tensor(A.indptr.astype(np.int32), **i32)
    cols = torch.as_tensor(A.indices.astype(np.int32), **i32)
    rows = torch.as_tensor(np.repeat(np.arange(A.shape[0], dtype=np.int32), np.diff(A.indptr)), **i32)
    return indptr, cols, rows


@dataclasses.dataclass
class LevelStructure:
    """Static structure of one level (and its transfer to the next coarser)."""

    n: int
    h: int
    w: int
    knots: tuple
    indptr: object
    cols: object
    rows: object
    regular: bool = False  # CSR rows are exactly the clipped 7 x 7 windows
    P: object = None  # host CSR (n x n_coarse)
    P1: tuple = None  # the 1-D factors of P (direction 0, direction 1), P = P1[1] (x) P1[0]
    radius: int = 0  # largest grid distance the level operator couples
    p_dev: tuple = None  # (indptr, cols, vals64, vals32)
    pt_dev: tuple = None
    t_pat: tuple = None  # pattern of P^T K (n_coarse x n): (indptr, cols, rows)

    @property
    def nnz(self):
        return int(self.cols.shape[0])


class Hierarchy:
    """Levels, transfers and Galerkin patterns of one scalar 2-D space and
    free set (cached on the space)."""

    def __init__(self, ctx, max_coarse=64, max_levels=16):
        space = ctx.space
        if space.dim != 2 or space.vec != 1:
            raise NotImplementedError("multigrid works on scalar 2-D tensor patches")
        if len(set(space.degree)) != 1:
            raise NotImplementedError("multigrid needs one degree in both directions")
        self.p = p = space.degree[0]
        dev = ctx.f2f_d.device
        lo, hi, fixed = free_rectangle(space.patch.n_cp_per_dir, ctx.dofmap.prescribed)
        knots = tuple(np.asarray(k, float) for k in space.patch.knot_arrays())
        pat = ctx.assembly
        h, w = hi[1] - lo[1] + 1, hi[0] - lo[0] + 1
        levels = [LevelStructure(pat.shape[0], h, w, knots, pat.indptr_d, pat.cols_d, pat.rows_d)]
        if h * w != pat.shape[0]:
            raise NotImplementedError("free dofs do not form the expected grid")
        K = sp.csr_matrix((np.ones(pat.nnz), pat.column_indices, pat.indptr), shape=pat.shape)
        while len(levels) < max_levels and levels[-1].n > max_coarse:
            fine = levels[-1]
            ck = tuple(coarsen_knots(t, p) for t in fine.knots)
            if any(c is None for c in ck):
                break
            P1 = []
            for d in range(2):
                T = change_of_basis(ck[d], p, fine.knots[d], p)  # fine coefs = T @ coarse coefs
                T[np.abs(T) < 1e-14] = 0.0
                nf, nc = T.shape
                rlo, rhi = (1 if fixed[(d, 0)] else 0), (nf - 2 if fixed[(d, 1)] else nf - 1)
                clo, chi = (1 if fixed[(d, 0)] else 0), (nc - 2 if fixed[(d, 1)] else nc - 1)
                P1.append(T[rlo: rhi + 1, clo: chi + 1])
            P = sp.csr_matrix(sp.kron(sp.csr_matrix(P1[1]), sp.csr_matrix(P1[0])))
            P.eliminate_zeros()
            P.sort_indices()
            PT = sp.csr_matrix(P.T)
            PT.sort_indices()
            Tpat = sp.csr_matrix(PT @ K)
            Tpat.sort_indices()
            Kc = sp.csr_matrix(Tpat @ P)
            Kc.data[:] = 1.0
            Kc.sort_indices()
            # the Galerkin kernel's box: every coarse row's fine stencil must fit
            # 16 x 16 fine grid points
            ptr = PT.indptr
            rows_f = PT.indices // fine.w
            cols_f = PT.indices % fine.w
            ext_r = np.maximum.reduceat(rows_f, ptr[:-1]) - np.minimum.reduceat(rows_f, ptr[:-1])
            ext_c = np.maximum.reduceat(cols_f, ptr[:-1]) - np.minimum.reduceat(cols_f, ptr[:-1])
            Kcoo = K.tocoo()
            radius = max(np.abs(Kcoo.col // fine.w - Kcoo.row // fine.w).max(),
                         np.abs(Kcoo.col % fine.w - Kcoo.row % fine.w).max())
            if max(ext_r.max(), ext_c.max()) + 1 + 2 * radius > 16 or radius > p:
                raise NotImplementedError("coarse-row stencil exceeds the Galerkin kernel's box")
            fine.P = P
            fine.P1 = (P1[0], P1[1])
            fine.radius = int(radius)
            fine.p_dev = self._transfer_dev(P, dev)
            fine.pt_dev = self._transfer_dev(PT, dev)
            fine.t_pat = _csr_dev(Tpat, dev)
            hc, wc = P1[1].shape[1], P1[0].shape[1]
            ip, cc, rr = _csr_dev(Kc, dev)
            levels.append(LevelStructure(Kc.shape[0], hc, wc, ck, ip, cc, rr,
                                         regular=regular_pattern(Kc.indptr, Kc.indices, hc, wc)))
            K = Kc
        Kcoo = K.tocoo()
        w_last = levels[-1].w
        levels[-1].radius = int(max(np.abs(Kcoo.col // w_last - Kcoo.row // w_last).max(initial=0),
                                    np.abs(Kcoo.col % w_last - Kcoo.row % w_last).max(initial=0)))
        self.levels = levels
        self._bands = {}

    @staticmethod
    def _transfer_dev(M, dev):
        import torch

        ip, cc, _ = _csr_dev(M, dev)
        v64 = torch.as_tensor(M.data, dtype=torch.float64, device=dev)
        return ip, cc, v64, v64.float()

    @property
    def n_levels(self):
        return len(self.levels)

    def tensor_bands(self):
        """Device coarse-column bands ``(f_lo, rw64)`` of every 1-D transfer
        (per level, per direction) for the tensor Galerkin passes, or None
        when an operator reaches beyond three grid points or a transfer's
        coupling does (a fine point of coarse column C, moved by up to three
        points, must stay in the support of columns within three of C)."""
        import torch

        if "tensor" in self._bands:
            return self._bands["tensor"]
        out = None
        if all(lv.radius <= GRID_RADIUS for lv in self.levels):
            dev = self.levels[0].indptr.device
            out = []
            for lv in self.levels[:-1]:
                per_dir = []
                for d in range(2):
                    tb = band_tables(lv.P1[d])
                    if tb is None or not tensor_coupling_ok(*tb):
                        return self._bands.setdefault("tensor", None)
                    c_lo, pw, f_lo, rw = tb
                    rw64 = np.zeros_like(rw, dtype=np.float64)
                    T = np.asarray(lv.P1[d], float)
                    for c in range(T.shape[1]):
                        k = min(KFC, T.shape[0] - f_lo[c])
                        rw64[c, :k] = T[f_lo[c]: f_lo[c] + k, c]
                    per_dir.append((torch.as_tensor(f_lo, device=dev), torch.as_tensor(rw64, device=dev)))
                out.append(per_dir)
        self._bands["tensor"] = out
        return out

    def multi_bands(self, depth=None):
        """Schedule and device bands of the multi-level Galerkin set-up: level
        m is formed from level ``src[m] = D floor((m - 1) / D)`` (D = ``depth``
        levels per launch, GAL_DEPTH by default) through the product of the
        1-D transfers in between (bands of GAL_KF[m - src] values per coarse
        column); ``(src, [(f_lo, rw64) per direction] per level, work
        doubles)``, or None when a product's band or coupling does not fit
        (the set-up then forms the levels one by one)."""
        import torch

        D = GAL_DEPTH if depth is None else depth
        if ("multi", D) in self._bands:
            return self._bands[("multi", D)]
        out = None
        lv = self.levels
        if self.tensor_bands() is not None:
            dev = lv[0].indptr.device
            src, bands, work = [0] * len(lv), [None] * len(lv), 0
            ok = 1 <= D <= 3
            for m in range(1, len(lv) if ok else 0):
                s = D * ((m - 1) // D)
                kf = GAL_KF[m - s]
                per_dir = []
                # wide bands are staged in LDS; the intermediates of a launch
                # are addressed with 32-bit byte offsets
                if (m - s > 1 and max(lv[m].h, lv[m].w) * (kf + 2 * GAL_PAD[kf]) > GAL_LDS_BAND) or \
                        3 * (2 * GRID_RADIUS + 1) ** 2 * lv[s].h * lv[m].w * 8 >= 2 ** 31:
                    ok = False
                    break
                for d in range(2):
                    T = np.asarray(lv[s].P1[d], float)
                    for l in range(s + 1, m):
                        T = T @ np.asarray(lv[l].P1[d], float)
                    T[np.abs(T) < 1e-14] = 0.0
                    # one level down: the narrow pass reads no padded rows
                    tb = wide_band(T, kf, pad=m - s > 1)
                    if tb is None:
                        ok = False
                        break
                    per_dir.append(tuple(torch.as_tensor(a, device=dev) for a in tb))
                if not ok:
                    break
                src[m], bands[m] = s, per_dir
            if ok:
                # the intermediates of a launch, and zeros past them for the
                # unclamped reads of the last one
                for s in range(0, len(lv) - 1, D):
                    work = max(work, sum((2 * GRID_RADIUS + 1) ** 2 * lv[s].h * lv[m].w
                                         for m in range(s + 1, min(s + D + 1, len(lv)))))
                work += GAL_TAIL * max(l.w for l in lv)
                out = (src, bands, work)
        self._bands[("multi", D)] = out
        return out

    def grid_bands(self, tail_level):
        """Device band tables of the transfers of levels ``0 .. tail_level-1``
        for the grid V-cycle, or None when a level's operator reaches beyond
        three grid points or a tile's transfer footprint exceeds the tile
        kernels' staging (see :func:`band_tables`)."""
        import torch

        if tail_level in self._bands:
            return self._bands[tail_level]
        out = None
        if 2 <= tail_level < self.n_levels and all(lv.radius <= GRID_RADIUS for lv in self.levels[:tail_level]):
            dev = self.levels[0].indptr.device
            bands = []
            for l in range(tail_level):
                per_dir = []
                for d in range(2):
                    tb = band_tables(self.levels[l].P1[d])
                    if tb is None or not band_spans_fit(*tb, level0_dir=d if l == 0 else None):
                        per_dir = None
                        break
                    per_dir.append(tuple(torch.as_tensor(a, device=dev) for a in tb))
                if per_dir is None:
                    bands = None
                    break
                bands.append(per_dir)
            out = bands
        self._bands[tail_level] = out
        return out


# the grid V-cycle's tile kernels: operator radius, band widths (coarse values
# per fine row and fine values per coarse column of a 1-D transfer), and the
# fine / coarse extents a tile stages: 16 x 16 tiles with a 3-point halo on
# the coarse levels; on level 0 row tiles of 64 (direction 0) x 4
# (direction 1) points whose prolongation stages 44 x 12 coarse points
GRID_RADIUS = 3
KPF, KFC = 4, 6
FINE_SPAN, COARSE_SPAN = 56, 24
L0_TILE = (64, 4)
L0_COARSE_SPAN = (44, 12)


def band_tables(T):
    """Band form of a 1-D transfer ``T`` (fine x coarse): ``c_lo[f]`` and
    ``pw[f, k] = T[f, c_lo[f] + k]`` (k < KPF) per fine row, ``f_lo[c]`` and
    ``rw[c, k] = T[f_lo[c] + k, c]`` (k < KFC) per coarse column; first
    indices non-decreasing. None if a row or column is wider than its band."""
    T = np.asarray(T, float)
    nf, nc = T.shape
    c_lo = np.zeros(nf, np.int32)
    pw = np.zeros((nf, KPF), np.float32)
    for f in range(nf):
        nz = np.flatnonzero(T[f])
        if len(nz):
            if nz[-1] - nz[0] >= KPF:
                return None
            c_lo[f] = nz[0]
            pw[f, : nz[-1] - nz[0] + 1] = T[f, nz[0]: nz[-1] + 1]
    f_lo = np.zeros(nc, np.int32)
    rw = np.zeros((nc, KFC), np.float32)
    for c in range(nc):
        nz = np.flatnonzero(T[:, c])
        if len(nz):
            if nz[-1] - nz[0] >= KFC:
                return None
            f_lo[c] = nz[0]
            rw[c, : nz[-1] - nz[0] + 1] = T[nz[0]: nz[-1] + 1, c]
    if np.any(np.diff(c_lo) < 0) or np.any(np.diff(f_lo) < 0):
        return None
    return c_lo, pw, f_lo, rw


# band widths of the product of one, two and three 1-D transfers (fine values
# per coarse column; dyadic cubic transfers need 5, 13 and 29), the zeros the
# passes stage around each column's band in LDS, the LDS they stage it in
# (doubles), and the zero values past the end of every FP64 stencil that the
# passes' unclamped reads may reach
GAL_KF = (0, KFC, 14, 30)
GAL_DEPTH = 3  # coarse levels formed per launch pair
GAL_PAD = {KFC: 9, 14: 16, 30: 28}
GAL_LDS_BAND = 4096
GAL_TAIL = 32


def wide_band(T, kf, pad=True):
    """Coarse-column band ``(f_lo, rw64)`` of a 1-D transfer (fine x coarse)
    with ``kf`` values per column, for the tensor Galerkin passes; None if a
    column is wider, first indices decrease, the coupling of the passes
    (every fine point of column C moved by up to GRID_RADIUS points lies only
    in columns within GRID_RADIUS of C) does not hold, or (``pad``) the
    neighbour offsets leave the padded LDS rows."""
    T = np.asarray(T, float)
    nf, nc = T.shape
    f_lo = np.zeros(nc, np.int32)
    rw = np.zeros((nc, kf), np.float64)
    for c in range(nc):
        nz = np.flatnonzero(T[:, c])
        if len(nz):
            if nz[-1] - nz[0] >= kf:
                return None
            f_lo[c] = nz[0]
        k = min(kf, nf - f_lo[c])
        rw[c, :k] = T[f_lo[c]: f_lo[c] + k, c]
    if np.any(np.diff(f_lo) < 0):
        return None
    # the close step reads KF + 6 consecutive padded band values of each
    # neighbour column from f_lo[C] - 3 - f_lo[C2]
    for dC in range(-GRID_RADIUS, GRID_RADIUS + 1) if pad else ():
        C = np.arange(max(0, -dC), min(nc, nc - dC))
        off = f_lo[C] - GRID_RADIUS - f_lo[C + dC]
        if len(off) and (off.min() < -GAL_PAD[kf] or off.max() > GAL_PAD[kf] - 2 * GRID_RADIUS):
            return None
    nzr = T != 0.0
    for C in range(nc):
        rows = np.flatnonzero(nzr[:, C])
        if not len(rows):
            continue
        g = np.arange(max(rows.min() - GRID_RADIUS, 0), min(rows.max() + GRID_RADIUS, nf - 1) + 1)
        cols = np.flatnonzero(nzr[g].any(axis=0))
        if len(cols) and np.abs(cols - C).max() > GRID_RADIUS:
            return None
    return f_lo, rw


def band_spans_fit(c_lo, pw, f_lo, rw, level0_dir=None, tile=16, halo=3):
    """Whether every tile's transfer footprint fits the tile kernels' staging:
    the fine points a coarse tile (with halo) restricts from, and the coarse
    points a fine tile (with halo) prolongs from; ``level0_dir`` (0 or 1) for
    the transfer out of level 0 along that direction (row tiles)."""
    nf, nc = len(c_lo), len(f_lo)
    for c0 in range(0, nc, tile):
        a, z = max(c0 - halo, 0), min(c0 + tile + halo - 1, nc - 1)
        if f_lo[z] + KFC - f_lo[a] > FINE_SPAN:
            return False
    ft, span = (tile, COARSE_SPAN) if level0_dir is None else (L0_TILE[level0_dir], L0_COARSE_SPAN[level0_dir])
    for f0 in range(0, nf, ft):
        a, z = max(f0 - halo, 0), min(f0 + ft + halo - 1, nf - 1)
        if c_lo[z] + KPF - c_lo[a] > span:
            return False
    return True


def regular_pattern(indptr, cols, h, w, radius=GRID_RADIUS):
    """Whether every CSR row (sorted columns) holds exactly the in-grid points
    of its (2 radius + 1)^2 window."""
    indptr = np.asarray(indptr)
    cols = np.asarray(cols)
    n = h * w
    if len(indptr) != n + 1:
        return False
    i = np.arange(n)
    r, c = i // w, i % w
    count = ((np.minimum(radius, h - 1 - r) - np.maximum(-radius, -r) + 1)
             * (np.minimum(radius, w - 1 - c) - np.maximum(-radius, -c) + 1))
    if not np.array_equal(np.diff(indptr), count):
        return False
    rows = np.repeat(i, count)
    dr, dc = cols // w - rows // w, cols % w - rows % w
    if np.abs(dr).max(initial=0) > radius or np.abs(dc).max(initial=0) > radius:
        return False
    key = rows.astype(np.int64) * n + cols
    return bool(np.all(np.diff(key) > 0))


def tensor_coupling_ok(c_lo, pw, f_lo, rw):
    """Whether the 1-D Galerkin pass of a transfer stays within the radius:
    every fine point f of coarse column C (P[f, C] != 0),
    moved by up to GRID_RADIUS points, lies only in columns C' with
    |C' - C| <= GRID_RADIUS."""
    nf, nc = len(c_lo), len(f_lo)
    for C in range(nc):
        for f in range(f_lo[C], min(f_lo[C] + KFC, nf)):
            if rw[C, f - f_lo[C]] == 0.0:
                continue
            for g in range(max(f - GRID_RADIUS, 0), min(f + GRID_RADIUS, nf - 1) + 1):
                cols = np.flatnonzero(pw[g]) + c_lo[g]
                if len(cols) and np.abs(cols - C).max() > GRID_RADIUS:
                    return False
    return True



# levels at or below this size take the dense tail (a 289-dof level: 0.3 MB
# FP32 with padded rows, one product per cycle instead of about a dozen
# launches; formed by batched cycles, whose cost grows with the square of the
# size); the dense product takes at most 2048 rows
TAIL_ROWS = 300
# the dense tail formed from the stencils: coarsest level and the level below
# the tail at most this large
DENSE_TAIL_COARSE, DENSE_TAIL_NEXT = 48, 128


def _free_key(dofmap):
    """Digest of a dof map's free set: contexts of one space with the same
    free set share one hierarchy."""
    return dofmap.key


def hierarchy(ctx, max_coarse=64):
    """Cached :class:`Hierarchy` of a method context's space and free set."""
    key = ("mg", _free_key(ctx.dofmap), max_coarse)
    cache = ctx.space._patterns
    if key not in cache:
        cache[key] = Hierarchy(ctx, max_coarse)
    return cache[key]


def galerkin(hier: Hierarchy, vals64, out=None, tbufs=None):
    """FP64 operator values of every level, ``K_{l+1} = P_l^T K_l P_l``, into
    ``out`` (per-level buffers, ``out[0]`` may be ``vals64`` itself) or new
    tensors; ``tbufs`` are optional buffers for the intermediate ``P^T K``."""
    import torch

    L = lib()
    st = stream()
    dev = vals64.device
    if out is None:
        out = [vals64] + [torch.empty(lv.nnz, dtype=torch.float64, device=dev) for lv in hier.levels[1:]]
    elif out[0] is not vals64:
        out[0].copy_(vals64)
    for l in range(hier.n_levels - 1):
        fine, coarse = hier.levels[l], hier.levels[l + 1]
        tp_ptr, tp_cols, tp_rows = fine.t_pat
        pt_ptr, pt_cols, pt_v64, _ = fine.pt_dev
        tvals = tbufs[l] if tbufs is not None else torch.empty(tp_cols.shape[0], dtype=torch.float64, device=dev)
        # T = P^T K (K symmetric: its CSR is the CSR of K^T)
        _lib.check(L.iga_masked_spgemm(tvals.shape[0], _lib.ptr(tp_rows), _lib.ptr(tp_cols),
                                       _lib.ptr(pt_ptr), _lib.ptr(pt_cols), _lib.ptr(pt_v64),
                                       _lib.ptr(fine.indptr), _lib.ptr(fine.cols), _lib.ptr(out[l]),
                                       _lib.ptr(tvals), st), "iga_masked_spgemm")
        # K_c = T P (B = P is given by the CSR of P^T)
        _lib.check(L.iga_masked_spgemm(coarse.nnz, _lib.ptr(coarse.rows), _lib.ptr(coarse.cols),
                                       _lib.ptr(tp_ptr), _lib.ptr(tp_cols), _lib.ptr(tvals),
                                       _lib.ptr(pt_ptr), _lib.ptr(pt_cols), _lib.ptr(pt_v64),
                                       _lib.ptr(out[l + 1]), st), "iga_masked_spgemm")
    return out


class MGOperator:
    """Numeric multigrid over persistent device buffers: Galerkin levels in FP64
    and FP32, diagonal scalings, the coarsest inverse and the smoothers
    (stencils per level, or damped Jacobi with ``omega`` where a level has
    none). :meth:`load` rewrites the buffers in place for a new operator, so
    the captured flexible-CG plans stay valid; one instance per hierarchy."""

    def __init__(self, hier: Hierarchy, omega=0.5, ks=7):
        import torch

        from jaxiga_amd import linear as LIN

        self.hier = hier
        self.ks = ks
        self.omega = float(omega)
        dev = hier.levels[0].indptr.device
        f64 = dict(dtype=torch.float64, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        self.patterns = [LIN.CSRPattern(lv.indptr, lv.cols, lv.rows, lv.n) for lv in hier.levels]
        self.vals64 = [torch.empty(lv.nnz, **f64) for lv in hier.levels]
        self.vals32 = [torch.empty(lv.nnz, **f32) for lv in hier.levels]
        self.dinv = [torch.empty(lv.n, **f64) for lv in hier.levels]
        self.dsq = [torch.empty(lv.n, **f32) for lv in hier.levels]
        c = hier.levels[-1]
        self._inv_work = torch.empty(2 * c.n * c.n, **f64)
        self.coarse_inv = torch.empty(c.n * c.n, **f32)
        self._stencil_buf = {}
        self.stencils = [None] * hier.n_levels
        # FP16 copy of level 0's stencil for the grid cycle (see set_stencils)
        self._stencil16 = None
        self._half = False
        n0 = hier.levels[0].n
        self.b = torch.empty(n0, **f64)
        self.x = torch.empty(n0, **f64)
        self._fwork = torch.empty(int(lib().iga_fcg_workspace(n0)), **f64)
        self._fwork32 = torch.empty(2 * n0, **f32)
        self._work = None
        self._struct = None
        self._setup = None
        self._plans = {}
        # the cycle below the first level with at most TAIL_ROWS dofs (level
        # >= 1, damped Jacobi) is formed as one dense matrix per operator
        self.tail_level = next((l for l in range(1, hier.n_levels - 1)
                                if hier.levels[l].n <= TAIL_ROWS), -1)
        if self.tail_level > 0:
            nt = hier.levels[self.tail_level].n
            self.tail = torch.empty(nt * int(lib().iga_mg_dense_ld(nt)), **f32)
            self._tail_work = None
        else:
            self.tail = None
        # grid form for the tile-kernel V-cycle: level operators above the
        # tail in stencil layout, transfers as 1-D bands
        self._bands = hier.grid_bands(self.tail_level) if self.tail_level > 0 else None
        self.grid_ops = None
        self._grid = None
        if self._bands is not None:
            nst = (2 * GRID_RADIUS + 1) ** 2
            self.grid_ops = [torch.empty(nst * hier.levels[l].n, **f32) for l in range(self.tail_level)]
            self.grid_op64 = torch.zeros(nst * n0 + GAL_TAIL, **f64)  # the flexible-CG operator
            # level 0's stencil divided by its diagonal, written with the layouts
            self.grid_feats = torch.empty(nst * n0, **f32)
            g = _lib.IgaMgGrid()
            g.n_levels = self.tail_level
            g.radius = GRID_RADIUS
            g.op64 = self.grid_op64.data_ptr()
            g.feats = self.grid_feats.data_ptr()
            for l, (op, bands) in enumerate(zip(self.grid_ops, self._bands)):
                g.op[l] = op.data_ptr()
                for d in range(2):
                    c_lo, pw, f_lo, rw = bands[d]
                    B = g.band[l][d]
                    B.c_lo, B.pw, B.f_lo, B.rw = c_lo.data_ptr(), pw.data_ptr(), f_lo.data_ptr(), rw.data_ptr()
            self._grid = g
        # tensor Galerkin set-up (with the grid cycle): FP64 stencils of every
        # level, the pass intermediate, and the FP32 stencils the cycle reads
        self._tensor = None
        tb = hier.tensor_bands() if self._grid is not None else None
        if tb is not None:
            nst = (2 * GRID_RADIUS + 1) ** 2
            lv = hier.levels
            # zeros past each FP64 stencil: the multi-level passes read up to
            # GAL_TAIL values beyond the last grid point
            self.grid_s64 = [self.grid_op64] + [torch.zeros(nst * l.n + GAL_TAIL, **f64) for l in lv[1:]]
            self._gal_work = torch.empty(nst * max(lv[l].h * lv[l + 1].w for l in range(len(lv) - 1)), **f64)
            sg = _lib.IgaMgSetupGrid()
            for l in range(len(lv)):
                sg.op64[l] = self.grid_s64[l].data_ptr()
                sg.op32[l] = self.grid_ops[l].data_ptr() if l < len(self.grid_ops) else None
            for l, per_dir in enumerate(tb):
                for d in range(2):
                    sg.f_lo[l][d] = per_dir[d][0].data_ptr()
                    sg.rw[l][d] = per_dir[d][1].data_ptr()
            sg.feats = self.grid_feats.data_ptr()
            sg.work = self._gal_work.data_ptr()
            # coarse levels up to three at a time from one finer level
            mb = hier.multi_bands()
            if mb is not None:
                src, bands, nwork = mb
                self._gal_multi_work = torch.zeros(nwork, **f64)
                sg.gal_multi = 1
                sg.gal_work = self._gal_multi_work.data_ptr()
                for m in range(1, len(lv)):
                    sg.gal_src[m] = src[m]
                    for d in range(2):
                        sg.gal_f_lo[m][d] = bands[m][d][0].data_ptr()
                        sg.gal_rw[m][d] = bands[m][d][1].data_ptr()
            for l in range(len(lv)):
                sg.regular[l] = 1 if lv[l].regular else 0
            # the dense tail straight from the stencils (small levels below it)
            t = self.tail_level
            if (t > 0 and lv[-1].n <= DENSE_TAIL_COARSE and lv[t + 1].n <= DENSE_TAIL_NEXT):
                sg.tail_level = t
                sg.omega = self.omega
                sg.tail = self.tail.data_ptr()
            self._tensor = sg

    def _setup_struct(self):
        import torch

        if getattr(self, "_setup", None) is None:
            h = self.hier
            m = _lib.IgaMgSetup()
            m.n_levels = h.n_levels
            m.radius = h.p
            for l, lv in enumerate(h.levels):
                L = m.level[l]
                L.n, L.h, L.w, L.nnz = lv.n, lv.h, lv.w, lv.nnz
                L.indptr, L.cols = lv.indptr.data_ptr(), lv.cols.data_ptr()
                L.vals64, L.vals32 = self.vals64[l].data_ptr(), self.vals32[l].data_ptr()
                L.dinv, L.dsq = self.dinv[l].data_ptr(), self.dsq[l].data_ptr()
                if lv.pt_dev is not None:
                    ip, cc, v64, _ = lv.pt_dev
                    L.pt_ptr, L.pt_cols, L.pt_vals = ip.data_ptr(), cc.data_ptr(), v64.data_ptr()
            m.coarse_inv = self.coarse_inv.data_ptr()
            m.inv_work = self._inv_work.data_ptr()
            if self._tensor is not None:
                m.grid = _lib.ctypes.pointer(self._tensor)
                if self._tensor.tail_level > 0:
                    nw = int(lib().iga_mg_tail_dense_doubles(m, self._tensor.tail_level))
                    self._tail_dense_work = torch.empty(nw, dtype=torch.float64, device=self.b.device)
                    self._tensor.tail_work = self._tail_dense_work.data_ptr()
            self._setup = m
        return self._setup

    def split_load(self):
        """Whether :meth:`load` can run in two stages (the tensor set-up)."""
        return self._tensor is not None

    def load(self, vals64, stencils=None, stage=0, vals_slot=None):
        """Write a new level-0 operator (FP64 CSR values) and its smoothers:
        one library call forms every Galerkin level, the FP32 copies, the
        diagonal scalings and the coarsest inverse, a second the dense matrix
        of the cycle below the tail level.

        With the tensor set-up, ``stage=1`` writes level 0 only (its stencil
        layouts, the generator features, copies and scalings) and ``stage=2``
        the rest, so work that reads the features can run on another stream
        beside stage 2; ``stage=0`` does both. Stage 2 may also run as
        ``stage=4`` (its first launch) and ``stage=8`` (the rest), stage 1 as
        ``stage=16`` (without level 0's CSR copies) and ``stage=32`` (the
        copies alone).
        ``vals_slot`` (tensor set-up): a one-element int64 device tensor
        holding the address of the values, which level 0 is then read
        through (``vals64`` is not read)."""
        if stage and not self.split_load():
            raise ValueError("a staged load needs the tensor set-up")
        if vals_slot is not None and self._tensor is None:
            raise ValueError("reading the values through a slot needs the tensor set-up")
        s = self._setup_struct()
        if self._tensor is not None:
            self._tensor.stages = stage
            self._tensor.vals0_slot = vals_slot.data_ptr() if vals_slot is not None else None
        try:
            _lib.check(lib().iga_mg_load(s, _lib.ptr(vals64), stream()), "iga_mg_load")
        finally:
            if self._tensor is not None:
                self._tensor.vals0_slot = None
        if stage in (1, 4, 16, 32):
            return self
        self.set_stencils(stencils if stencils is not None else [None] * self.hier.n_levels)
        self._form_tail()
        if self._grid is not None and self._tensor is None:
            # the tensor set-up writes the stencil layouts itself
            _lib.check(lib().iga_mg_grid_ops(self.struct(), _lib.ptr(self.vals64[0]), stream()),
                       "iga_mg_grid_ops")
        return self

    def grid_active(self):
        """Whether the V-cycle from level 0 runs as the grid tile kernels: a
        grid form exists, the dense tail applies and only level 0 may carry a
        (7 x 7) stencil."""
        return (self._grid is not None and self.tail_active() and self.ks == 2 * GRID_RADIUS + 1
                and all(s is None for s in self.stencils[1:]))

    def _form_tail(self):
        import torch

        if self.tail is None or not self.tail_active():
            return
        if self._tensor is not None and self._tensor.tail_level > 0:
            return  # formed by iga_mg_load from the stencils
        m = self.struct()
        if self._tail_work is None:
            nw = int(lib().iga_mg_tail_floats(m, self.tail_level))
            self._tail_work = torch.empty(nw, dtype=torch.float32, device=self.b.device)
        _lib.check(lib().iga_mg_tail(m, self.tail_level, _lib.ptr(self._tail_work), _lib.ptr(self.tail),
                                     stream()), "iga_mg_tail")

    def tail_active(self):
        """The dense tail applies while no level at or below it has a stencil."""
        return self.tail is not None and all(s is None for s in self.stencils[self.tail_level:])

    def stencil_buffer(self, l):
        """The persistent stencil buffer of level l (``ks*ks*n_l`` floats); a
        stencil written there is taken by :meth:`set_stencils` without a copy."""
        import torch

        buf = self._stencil_buf.get(l)
        if buf is None:
            buf = torch.empty(self.ks * self.ks * self.hier.levels[l].n, dtype=torch.float32,
                              device=self.b.device)
            self._stencil_buf[l] = buf
        return buf

    def write_half(self, stencil0):
        """Write the FP16 copy of level 0's stencil (on the current stream)
        for the grid cycle; returns the ``half`` argument that hands it to
        :meth:`set_stencils` ("ready"), or None where the grid cycle does not
        apply."""
        import torch

        if stencil0 is None or self._grid is None:
            return None
        if self._stencil16 is None:
            self._stencil16 = torch.empty(stencil0.numel(), dtype=torch.float16, device=stencil0.device)
        self._stencil16.copy_(stencil0.reshape(-1))
        return "ready"

    def set_stencils(self, stencils, half=None):
        """Smoother stencils per level (None: damped Jacobi), copied into the
        persistent buffers unless they are those buffers. ``half``: the grid
        cycle's level-0 smoothers read an FP16 copy of level 0's stencil --
        "copy" writes it (on the stream), "ready" takes the one written last
        (by a captured set-up that made it); None: FP32 only."""
        import torch

        new = []
        for l, s in enumerate(list(stencils) + [None] * (self.hier.n_levels - len(stencils))):
            if s is None:
                new.append(None)
                continue
            buf = self.stencil_buffer(l)
            if s.data_ptr() != buf.data_ptr():
                buf.copy_(s.reshape(-1))
            new.append(buf)
        if [x is None for x in new] != [x is None for x in self.stencils]:
            self._struct = None
        half_on = (half == "copy" or (half == "ready" and self._stencil16 is not None)) and \
            new[0] is not None and self._grid is not None
        if half_on and half == "copy":
            if self._stencil16 is None:
                self._stencil16 = torch.empty_like(new[0], dtype=torch.float16)
            self._stencil16.copy_(new[0])
        if half_on != self._half:
            self._struct = None
        self._half = half_on
        was = self.tail_active()
        self.stencils = new
        if self.tail_active() and not was:
            self._form_tail()

    def struct(self):
        import torch

        if self._struct is not None:
            return self._struct
        # one descriptor per pattern of levels with stencils (their buffers
        # are persistent), built once
        key = (tuple(s is None for s in self.stencils), self.tail_active(), self._half)
        cached = self.__dict__.setdefault("_structs", {}).get(key)
        if cached is not None:
            self._struct = cached
            return cached
        h = self.hier
        m = _lib.IgaMg()
        m.n_levels = h.n_levels
        m.ks = self.ks
        for l, lv in enumerate(h.levels):
            L = m.level[l]
            L.n, L.h, L.w, L.nnz = lv.n, lv.h, lv.w, lv.nnz
            L.indptr, L.cols = lv.indptr.data_ptr(), lv.cols.data_ptr()
            L.vals, L.dsq = self.vals32[l].data_ptr(), self.dsq[l].data_ptr()
            L.stencil = self.stencils[l].data_ptr() if self.stencils[l] is not None else None
            if l == 0 and self._half:
                L.stencil16 = self._stencil16.data_ptr()
            L.omega = self.omega
            if lv.p_dev is not None:
                ip, cc, _, v32 = lv.p_dev
                L.p_nnz, L.p_ptr, L.p_cols, L.p_vals = cc.shape[0], ip.data_ptr(), cc.data_ptr(), v32.data_ptr()
                ip, cc, _, v32 = lv.pt_dev
                L.pt_nnz, L.pt_ptr, L.pt_cols, L.pt_vals = cc.shape[0], ip.data_ptr(), cc.data_ptr(), v32.data_ptr()
        m.coarse_inv = self.coarse_inv.data_ptr()
        if self.tail_active():
            m.tail_level = self.tail_level
            m.tail = self.tail.data_ptr()
        if self._grid is not None:
            m.grid = _lib.ctypes.pointer(self._grid)
        if self._work is None:
            nw = int(lib().iga_mg_work_floats(m))
            self._work = torch.empty(max(nw, 1), dtype=torch.float32, device=self.b.device)
        m.work = self._work.data_ptr()
        self._struct = m
        self._structs[key] = m
        return m

    def vcycle(self, b32, level=0, out=None):
        import torch

        x = torch.empty_like(b32) if out is None else out
        _lib.check(lib().iga_mg_vcycle(self.struct(), level, _lib.ptr(b32), _lib.ptr(x), stream()),
                   "iga_mg_vcycle")
        return x

    def _plan(self, graph_iters):
        key = (tuple(s is None for s in self.stencils), self.tail_active(), self._half, graph_iters)
        plan = self._plans.get(key)
        if plan is None:
            import torch

            lv = self.hier.levels[0]
            handle = _lib.c_void_p()
            cur = torch.cuda.current_stream()
            side = side_stream(self.b.device, "fcg") if graph_iters > 0 else cur
            side.wait_stream(cur)
            with torch.cuda.stream(side):  # graph capture needs a non-default stream
                _lib.check(lib().iga_fcg_mg_plan(self.struct(), lv.n, lv.nnz, _lib.ptr(lv.indptr),
                                                 _lib.ptr(lv.cols), _lib.ptr(self.vals64[0]),
                                                 _lib.ptr(self.b), _lib.ptr(self.x), _lib.ptr(self._fwork),
                                                 _lib.ptr(self._fwork32), int(graph_iters), stream(),
                                                 _lib.ctypes.byref(handle)),
                           "iga_fcg_mg_plan")
            cur.wait_stream(side)
            plan = _Plan(handle)
            self._plans[key] = plan
        return plan

    def fcg(self, b, tol=1e-10, maxiter=1000, graph_iters=8, true_residual=False):
        """Flexible CG on the level-0 FP64 operator with one V-cycle per
        iteration; returns ``(x, iterations, rr, bb)`` (x a new tensor), and
        ``||b - K x||`` recomputed in FP64 from x as a fifth item when
        ``true_residual``."""
        import torch

        it, rr, bb, res = _lib.c_int(0), _lib.c_double(0.0), _lib.c_double(0.0), _lib.c_double(-1.0)
        plan = self._plan(graph_iters)  # captured once; replayed on the current stream
        if b.dtype != torch.float64 or b.device != self.b.device or b.numel() != self.b.numel():
            raise ValueError("fcg: b must be a float64 vector of the level-0 size on the operator's device")
        b = b.contiguous()
        x = torch.empty_like(self.x)  # filled by the run, behind the iterations
        # the run's first kernel reads b and keeps it in self.b
        _lib.check(lib().iga_fcg_mg_run(plan.handle, float(tol), int(maxiter), _lib.ctypes.byref(it),
                                        _lib.ctypes.byref(rr), _lib.ctypes.byref(bb),
                                        _lib.ctypes.byref(res) if true_residual else None,
                                        _lib.ptr(x), _lib.ptr(b), stream()),
                   "iga_fcg_mg_run")
        if true_residual:
            return x, it.value, rr.value, bb.value, res.value
        return x, it.value, rr.value, bb.value


class _Plan:
    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        try:
            lib().iga_fcg_mg_destroy(self.handle)
        except Exception:
            pass


def operator(hier: Hierarchy, omega=0.5, ks=7):
    """The persistent :class:`MGOperator` of a hierarchy."""
    key = (omega, ks)
    ops = hier.__dict__.setdefault("_operators", {})
    if key not in ops:
        ops[key] = MGOperator(hier, omega, ks)
    return ops[key]
