"""Batched assembly and solve of many uneven Poisson instances.

The reference batches by ``jax.vmap`` over parameters on one fixed space
(``tests/test_advanced.py:337-344``), with host callbacks running sequentially
per batch member (``README.md:117``). Here a batch of instances with different
degrees, mesh sizes and coefficient fields is grouped by shape (degree, mesh):

* per shape, one element-pass launch and one gather launch assemble every
  instance of the group against the shared CSR pattern (``iga_assemble`` with
  ``batch > 1``), the diagonal coming out of the same pass;
* the batch is solved by FP64 Jacobi-PCG to ``tol`` in one of two ways:
  ``segmented`` (``iga_batched_cg_seg``, the default): every instance's rows
  are cut into blocks of ``ROWS_PER_BLOCK`` rows and each iteration is three
  launches over the blocks of all instances, so the large instances stream
  through the whole GPU; or ``persistent`` (``iga_batched_cg``): one
  workgroup per instance runs its whole solve in one launch, largest first;
* the method and its lanes per row (and threads per workgroup) are chosen by
  a small autotuner that times the candidates on the batch itself;
* the batch pipeline with the persistent solve can be captured in a hipGraph
  (via a torch CUDA graph on a side stream) and replayed; the segmented solve
  reads an active-instance count back between groups of iterations.
"""

from __future__ import annotations

import dataclasses
import itertools
import time

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import methods as M
from jaxiga_amd.runtime import lib, stream


@dataclasses.dataclass
class InstanceSpec:
    degree: int
    n_el: int  # elements per side
    ell: float
    seed: int


def cfg4_specs(n_instances=512, seed=20260902, sigma=1.0):
    """512 unit-square Poisson instances: p in {2,3}, n in {16,32,64,128},
    GRF length scale in [0.05, 0.2] (all seeded)."""
    rng = np.random.default_rng(seed)
    p = rng.choice([2, 3], size=n_instances)
    n = rng.choice([16, 32, 64, 128], size=n_instances)
    ell = rng.uniform(0.05, 0.2, size=n_instances)
    return [InstanceSpec(int(a), int(b), float(c), seed * 1000 + i)
            for i, (a, b, c) in enumerate(zip(p, n, ell))]


class _Group:
    def __init__(self, degree, n_el, members, specs, sigma, device):
        import torch

        import jaxiga_amd as jx
        from jaxiga_amd import grf

        refine = int(round(np.log2(n_el)))
        V = jx.FunctionSpace(jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]])
                             .elevate(degree).refine(refine))
        prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))],
                          source=1.0)
        self.members = members
        self.degree, self.n_el = degree, n_el
        self.ctx = M.build_context(prob)
        self.pattern = self.ctx.assembly
        B = len(members)
        n_e, n_q = V.n_elems, self.ctx.n_q1 ** 2
        xq = self.ctx.gauss_points()
        coef = torch.empty((B, n_e, n_q), dtype=torch.float64, device=device)
        for k, i in enumerate(members):
            om, ph = grf.rff_params(specs[i].seed, 2, specs[i].ell)
            coef[k] = torch.exp(grf.rff_eval(xq, om, ph, sigma))
        self.coef = coef
        nf, nnz = self.ctx.n_free, self.pattern.nnz
        f64 = dict(dtype=torch.float64, device=device)
        self.vals = torch.empty((B, nnz), **f64)
        self.rhs = torch.zeros((B, nf), **f64)
        self.diag = torch.zeros((B, nf), **f64)
        self.dinv = torch.empty((B, nf), **f64)
        self.x = torch.empty((B, nf), **f64)
        self.work = torch.empty((B, 4 * nf), **f64)
        nb = int(lib().iga_assemble_workspace(V.device().struct, _lib.IGA_POISSON, _lib.IGA_FP64))
        self.ke = torch.empty(nb * B, dtype=torch.uint8, device=device)
        args = _lib.IgaAsmArgs()
        args.problem, args.vec, args.precision = _lib.IGA_POISSON, 1, _lib.IGA_FP64
        args.coef_kind, args.coef_points = _lib.IGA_COEF_POINTS, coef.data_ptr()
        args.src_kind = _lib.IGA_SRC_CONST
        args.src_const[0] = 1.0
        args.full_to_free, args.freed = self.ctx.f2f_d.data_ptr(), self.ctx.free_d.data_ptr()
        args.gfull = self.ctx.gfull_d.data_ptr()
        args.nnz, args.rows, args.cols = nnz, self.pattern.rows_d.data_ptr(), self.pattern.cols_d.data_ptr()
        args.indptr = self.pattern.indptr_d.data_ptr()
        args.vals, args.ke_work, args.ke_work_bytes = self.vals.data_ptr(), self.ke.data_ptr(), self.ke.numel()
        args.rhs, args.diag = self.rhs.data_ptr(), self.diag.data_ptr()
        args.batch, args.n_free = B, nf
        self.args = args

    def assemble(self):
        self.rhs.zero_()
        self.diag.zero_()
        _lib.check(lib().iga_assemble(self.ctx.space.device().struct, self.ctx.rule(), self.args,
                                      stream()), "iga_assemble")
        _lib.check(lib().iga_inv_guard(self.diag.numel(), _lib.ptr(self.diag), _lib.ptr(self.dinv),
                                       stream()), "iga_inv_guard")


ROWS_PER_BLOCK = 256

INST_DTYPE = np.dtype([("n", "<i4"), ("pad", "<i4"), ("indptr", "<u8"), ("cols", "<u8"),
                       ("vals", "<u8"), ("dinv", "<u8"), ("b", "<u8"), ("x", "<u8"), ("work", "<u8")])


class BatchedPoisson:
    """A batch of uneven unit-square Poisson instances assembled and solved together."""

    def __init__(self, specs, sigma=1.0, device="cuda", rows_per_block=ROWS_PER_BLOCK):
        import torch

        self.specs = specs
        shapes = {}
        for i, s in enumerate(specs):
            shapes.setdefault((s.degree, s.n_el), []).append(i)
        self.groups = [_Group(p, n, mem, specs, sigma, device) for (p, n), mem in sorted(shapes.items())]
        rec = np.zeros(len(specs), INST_DTYPE)
        sizes = np.zeros(len(specs), np.int64)
        for g in self.groups:
            nf = g.ctx.n_free
            for k, i in enumerate(g.members):
                rec[i] = (nf, 0, g.pattern.indptr_d.data_ptr(), g.pattern.cols_d.data_ptr(),
                          g.vals[k].data_ptr(), g.dinv[k].data_ptr(), g.rhs[k].data_ptr(),
                          g.x[k].data_ptr(), g.work[k].data_ptr())
                sizes[i] = g.pattern.nnz
        self.insts = torch.as_tensor(rec.view(np.uint8), device=device)
        self.order = torch.as_tensor(np.argsort(-sizes, kind="stable").astype(np.int32), device=device)
        self.iters = torch.zeros(len(specs), dtype=torch.int32, device=device)
        self.rel = torch.zeros(len(specs), dtype=torch.float64, device=device)
        # segmented solve: block map and workspace
        n_rows = np.ascontiguousarray(rec["n"], dtype=np.int32)
        L = lib()
        self.rows_per_block = int(rows_per_block)
        nb = int(L.iga_batched_cg_seg_plan(len(specs), n_rows.ctypes.data, self.rows_per_block, None, None,
                                           None))
        if nb < 0:
            raise RuntimeError(f"iga_batched_cg_seg_plan failed ({nb})")
        blk_inst, blk_row0 = np.empty(nb, np.int32), np.empty(nb, np.int32)
        inst_blk0 = np.empty(len(specs) + 1, np.int32)
        L.iga_batched_cg_seg_plan(len(specs), n_rows.ctypes.data, self.rows_per_block, blk_inst.ctypes.data,
                                  blk_row0.ctypes.data, inst_blk0.ctypes.data)
        self.n_blocks = nb
        self.seg_map = [torch.as_tensor(a, device=device) for a in (blk_inst, blk_row0, inst_blk0)]
        self.seg_work = torch.empty(int(L.iga_batched_cg_seg_workspace(len(specs), nb)), dtype=torch.uint8,
                                    device=device)
        self.n_rows = n_rows
        self.nnz = sizes
        self.method, self.lpr, self.threads = "segmented", 8, 256
        self.n_elements = sum(g.ctx.space.n_elems * len(g.members) for g in self.groups)

    def assemble(self):
        for g in self.groups:
            g.assemble()

    def solve(self, tol=1e-10, maxiter=20000, lpr=None, threads=None, method=None):
        method = method or self.method
        if method == "segmented":
            m = self.seg_map
            _lib.check(lib().iga_batched_cg_seg(
                len(self.specs), _lib.ptr(self.insts), self.n_blocks, _lib.ptr(m[0]), _lib.ptr(m[1]),
                _lib.ptr(m[2]), self.rows_per_block, lpr or self.lpr, float(tol), int(maxiter),
                _lib.ptr(self.seg_work), self.seg_work.numel(), _lib.ptr(self.iters), _lib.ptr(self.rel),
                stream()), "iga_batched_cg_seg")
        elif method == "persistent":
            _lib.check(lib().iga_batched_cg(len(self.specs), _lib.ptr(self.insts), _lib.ptr(self.order),
                                            lpr or self.lpr, threads or self.threads, float(tol),
                                            int(maxiter), _lib.ptr(self.iters), _lib.ptr(self.rel),
                                            stream()), "iga_batched_cg")
        else:
            raise ValueError(f"unknown batched solve method {method!r}")

    def run(self, tol=1e-10, method=None):
        self.assemble()
        self.solve(tol, method=method)

    def autotune(self, candidates=None, probe_iters=40):
        """Pick the solve method and its lanes per row (and, for the
        persistent solve, threads per workgroup) by timing ``probe_iters``
        iterations of every candidate on this batch."""
        import torch

        if candidates is None:
            candidates = [("segmented", lpr, 256) for lpr in (8, 16, 32)] + [
                ("persistent", lpr, thr) for lpr, thr in itertools.product((8, 16, 32), (256, 512, 1024))]
        self.assemble()
        best, timings = None, {}
        for method, lpr, thr in candidates:
            self.solve(1e-30, probe_iters, lpr, thr, method)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.solve(1e-30, probe_iters, lpr, thr, method)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            timings[(method, lpr, thr)] = dt
            if best is None or dt < timings[best]:
                best = (method, lpr, thr)
        self.method, self.lpr, self.threads = best
        return best, timings

    def capture(self, tol=1e-10):
        """A replayable graph of the whole batch (assembly + the persistent
        solve, which needs no host synchronisation)."""
        import torch

        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run(tol, method="persistent")  # warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self.run(tol, method="persistent")
        return g

    def results(self):
        return self.iters.cpu().numpy(), self.rel.cpu().numpy()

    def instance(self, i):
        """(group, index in group) of instance i."""
        for g in self.groups:
            if i in g.members:
                return g, g.members.index(i)
        raise KeyError(i)
