#!/usr/bin/env python
"""Benchmark of the Galerkin hot path on one MI355X.

Workload (N=1): 2-D plane-strain linear elasticity on the quarter annulus
r in [1, 2] (rational quadratic in the angular direction), degree elevated to
p=3 and refined to 256x256 Bezier elements (65 536 elements, 134 162 dofs,
12 923 976 free-free nonzeros), E=1, nu=0.3, clamped inner edge, unit normal
traction on the outer edge.

A step is one full assembly of that operator with its inputs already in HBM:
basis evaluation at the 16 Gauss points of every element, element stiffness,
scatter-add into the CSR values of the free-free block, and the load vector
(Neumann work and Dirichlet lifting). ``value`` is elements assembled per
second (Melem/s). The line also carries the FP64 Jacobi-CG time to a relative
residual of 1e-10 on the same operator, the roofline of the assembly kernel
(timed with HIP events on the stream it runs on), and the CPU oracle timed on
a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
REFINE = 8  # 2**8 = 256 elements per direction


def build_problem(refine=REFINE):
    import jaxiga_amd as jx

    patch = jx.primitives.quarter_annulus(1.0, 2.0).elevate(3).refine(refine)
    V = jx.FunctionSpace(patch, vec=2)
    prob = jx.LinearElasticity(V, plane="strain", dirichlet=[jx.DirichletBC(0.0, where="inner")],
                               neumann=[jx.Neumann(lambda x, n, p: n, where="outer")])
    return V, prob, {"E": 1.0, "nu": 0.3}


def measured_traffic(path=os.path.join(ROOT, "profiles", "assembly_elasticity_256.md")):
    """HBM bytes per launch of the assembly kernels from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE runs, scaled by a device copy), keyed
    by kernel name; empty if the summary is absent."""
    out = {}
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    for item in line.split()[1:]:
                        k, v = item.split("=")
                        out[k] = int(v)
    except OSError:
        pass
    return out


def elasticity_flops_per_element(n_q=16, n_s=3, n_lv=32):
    # element contraction K_e = sum_q w B^T C B (basis evaluation excluded)
    return 2 * n_q * (n_s * n_s * n_lv + n_s * n_lv * n_lv)


def hbm_roofline(kernel, nbytes, ms, algorithmic, traffic=None):
    """Roofline entry of a bandwidth-bound kernel: algorithmic bytes per launch
    over its measured duration, against the HBM peak."""
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "kernel": kernel,
            "kernel_ms": ms, "algorithmic": algorithmic, "frac": achieved / PEAK_HBM_GBS, "traffic": traffic}


def graph_ms(launch, per=20, reps=5):
    """Average device time (ms) of one launch for calls too short for eager
    timing (the host's enqueue time per call would exceed the kernel's):
    ``per`` calls captured once in a graph, the replay timed between HIP
    events on the current stream, the median over ``reps`` replays."""
    import torch

    launch()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        for _ in range(per):
            launch()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g.replay()
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / per)
    del g
    return statistics.median(times)


def event_ms(launch, batches=5, per=10):
    """Average duration (ms) of one launch: batches of ``per`` back-to-back
    launches on the current stream between two HIP events, each batch
    started behind an untimed launch (so the GPU is already busy when the
    first event is recorded); the median over batches, robust to a host
    pause inside one batch."""
    import torch

    times = []
    for _ in range(batches):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        launch()
        e0.record()
        for _ in range(per):
            launch()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / per)
    return statistics.median(times)


def cpu_baseline(sample_refine=7):
    """Oracle FP64 assembly (basis, K_e, CSR scatter-add) on the same geometry
    at 2**sample_refine elements per direction; median of 3 after a warm-up."""
    from oracle import assembly as OA
    from oracle import forms as OF
    from oracle import geometry as OG
    from oracle import pipeline as OP
    from oracle import space as OS

    VO = OS.function_space(OG.quarter_annulus(1.0, 2.0).elevate(3).refine(sample_refine), vec=2)
    params = {"E": 1.0, "nu": 0.3}
    presc, vals = OF.resolve_dirichlet(VO, [OF.DirichletBC(0.0, "inner")], params)
    dm = OS.make_dofmap(VO.n_dofs, presc, vals)
    pat = OA.csr_pattern(VO.elem_dofs, VO.n_dofs, VO.vec, dm.free)

    def once():
        basis = OS.gauss_basis(VO)
        Ke = OP.element_tangents("elasticity", VO, basis, params, "strain")
        return OA.assemble_csr(pat, Ke)

    once()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        once()
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {"value": VO.n_elems / t / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 assembly (basis + K_e + CSR scatter-add) of the same quarter-annulus "
                      f"elasticity operator at {2**sample_refine}x{2**sample_refine} elements, "
                      f"median of 3 runs of {t:.2f} s"}


def learned_pc_bench(refine=8, n_train=88, steps=300, batch=4, test_seeds=tuple(range(10_000, 10_032)),
                     reps=3, log=print, cpu=True):
    """Time to a 1e-10 relative residual on the 256x256 cubic variable-coefficient
    Poisson problem (a = exp(GRF), l = 0.1, sigma = 1.5, smooth random source):
    FP64 Jacobi-CG versus flexible CG with the learned multigrid preconditioner
    (FP32 V-cycle), on coefficient seeds the generator never saw. The generator
    is trained first, on the GPU, from pairs the solver generates on
    ``n_train`` training seeds (training time reported, not counted)."""
    import torch

    import jaxiga_amd as jx
    from jaxiga_amd import grf
    from jaxiga_amd import learned as LN
    from jaxiga_amd import linear as LIN

    V = jx.FunctionSpace(jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(3).refine(refine))
    t0 = time.perf_counter()
    data = LN.TrainingSet()
    for seed in range(n_train):
        ctx, vals, rhs = grf.poisson_instance(refine, seed, space=V)
        data.add(ctx, vals, rhs)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    n_pairs = data.n_pairs
    log(f"[bench] learned PC: {n_pairs} residual/error pairs from {n_train} fields in {t_data:.1f} s")
    gen = LN.generator(3, seed=1)
    t0 = time.perf_counter()
    hist = LN.train(gen, data, steps=steps, lr=1e-3, batch=batch)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"[bench] trained {steps} steps in {t_train:.1f} s, loss {hist[0]:.3f} -> {np.mean(hist[-10:]):.3f}")
    del data
    rows = []
    cpu_out = None

    def timed_cg(pattern, vals, rhs, work, single):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = LIN.cg_csr(pattern, vals, rhs, 1e-10, workspace=work, single_reduction=single)
        torch.cuda.synchronize()
        return res, time.perf_counter() - t0

    # warm-up on an operator outside the training and test seeds: first-use
    # captures (set-up graph, iteration graphs) and code loading stay out of
    # the timings; every timed operator below is then met for the first time
    ctx, vals, rhs = grf.poisson_instance(refine, 20_000, space=V)
    for _ in range(2):
        LN.solve(ctx, vals, rhs, gen)
        LN.solve(ctx, vals, rhs, None)
        LIN.cg_csr(ctx.assembly, vals, rhs, 1e-10, single_reduction=True)
        LIN.cg_csr(ctx.assembly, vals, rhs, 1e-10)
    torch.cuda.synchronize()
    for k, seed in enumerate(test_seeds):
        ctx, vals, rhs = grf.poisson_instance(refine, seed, space=V)
        pattern = ctx.assembly
        work = torch.empty(int(_lib_mod().iga_cg_workspace(ctx.n_free)), dtype=torch.float64, device="cuda")
        # first contact with the new operator: the time-to-solution a caller
        # sees. The two methods alternate which one meets the operator first
        # (the second finds its values in the caches).
        if k % 2 == 0:
            first_jac, t_jac_first = timed_cg(pattern, vals, rhs, work, True)
            first_le = LN.solve(ctx, vals, rhs, gen)
        else:
            first_le = LN.solve(ctx, vals, rhs, gen)
            first_jac, t_jac_first = timed_cg(pattern, vals, rhs, work, True)
        t_first = first_le.seconds
        LN.solve(ctx, vals, rhs, None)
        LIN.cg_csr(pattern, vals, rhs, 1e-10, workspace=work)
        tj, tj1, tu, tl = [], [], [], []
        for _ in range(reps):
            jac, t = timed_cg(pattern, vals, rhs, work, False)
            tj.append(t)
            jac1, t = timed_cg(pattern, vals, rhs, work, True)
            tj1.append(t)
            un = LN.solve(ctx, vals, rhs, None)
            tu.append(un.seconds)
            le = LN.solve(ctx, vals, rhs, gen)
            tl.append(le.seconds)
        rn, bn = LIN.residual_norms(pattern, vals, jac.x, rhs)
        rn1, _ = LIN.residual_norms(pattern, vals, first_jac.x, rhs)
        diff = float((le.x - jac.x).abs().max() / jac.x.abs().max())
        t_base = min(statistics.median(tj), statistics.median(tj1))
        row = {"seed": seed,
               "first_contact_ms": {"jacobi_cg1_fp64": [first_jac.iterations, t_jac_first * 1e3],
                                    "learned_mg_fcg": [first_le.iterations, t_first * 1e3,
                                                       first_le.setup_seconds * 1e3],
                                    "order": "jacobi first" if k % 2 == 0 else "learned first"},
               "jacobi_cg_fp64": [jac.iterations, statistics.median(tj) * 1e3],
               "jacobi_cg1_fp64": [jac1.iterations, statistics.median(tj1) * 1e3],
               "mg_fcg_untrained": [un.iterations, statistics.median(tu) * 1e3],
               "learned_mg_fcg": [le.iterations, statistics.median(tl) * 1e3, le.setup_seconds * 1e3,
                                  le.fallback_iterations],
               "relative_residual": [rn / bn, rn1 / bn, un.relative_residual, le.relative_residual,
                                     first_le.relative_residual],
               "max_rel_diff_vs_jacobi_cg": diff}
        row["speedup_first_contact"] = t_jac_first / t_first
        row["speedup"] = t_base / statistics.median(tl)
        log(f"[bench] seed {seed}: first contact: jacobi-cg1 {first_jac.iterations} it {t_jac_first * 1e3:.2f} ms, "
            f"learned {first_le.iterations} it {t_first * 1e3:.2f} ms -> {row['speedup_first_contact']:.2f}x | "
            f"repeated: jacobi-cg {jac.iterations} it {statistics.median(tj) * 1e3:.2f} ms, "
            f"cg1 {jac1.iterations} it {statistics.median(tj1) * 1e3:.2f} ms, "
            f"untrained {un.iterations} it {statistics.median(tu) * 1e3:.2f} ms, "
            f"learned {le.iterations} it {statistics.median(tl) * 1e3:.2f} ms "
            f"(setup {le.setup_seconds * 1e3:.2f} ms) -> {row['speedup']:.2f}x")
        rows.append(row)
        if cpu and len(rows) == 1:
            cpu_out = cpu_time_to_solve(pattern, vals, rhs)
            log(f"[bench] CPU time to solve (seed {seed}): SciPy Jacobi-CG "
                f"{cpu_out['scipy_jacobi_cg']['seconds']:.2f} s ({cpu_out['scipy_jacobi_cg']['iterations']} it), "
                f"SuperLU {cpu_out['superlu']['seconds']:.2f} s")
    return {"workload": f"unit-square Poisson, p=3, {2**refine}x{2**refine} elements, a=exp(GRF l=0.1 sigma=1.5), "
                        f"smooth random source, rel. residual 1e-10, {len(rows)} unseen coefficient seeds "
                        f"{test_seeds[0]}..{test_seeds[-1]}",
            "per_seed_columns": {"first_contact_ms": "first solve of the unseen operator by each method "
                                                     "([iterations, ms] / [iterations, ms, setup_ms])",
                                 "jacobi_cg_fp64": ["iterations", "ms (median of repeats, textbook CG)"],
                                 "jacobi_cg1_fp64": ["iterations", "ms (median of repeats, single-reduction CG)"],
                                 "mg_fcg_untrained": ["iterations", "ms"],
                                 "learned_mg_fcg": ["iterations", "ms", "setup_ms", "fallback_iterations"],
                                 "relative_residual": ["jacobi_cg", "jacobi_cg1 first", "untrained", "learned",
                                                       "learned first"]},
            "training": {"fields": n_train, "pairs": n_pairs, "steps": steps, "seconds": t_train,
                         "data_seconds": t_data, "loss_first": hist[0], "loss_last10": float(np.mean(hist[-10:]))},
            "per_seed": rows,
            "speedup_first_contact_median": statistics.median(r["speedup_first_contact"] for r in rows),
            "speedup_first_contact_min": min(r["speedup_first_contact"] for r in rows),
            "speedup_first_contact_max": max(r["speedup_first_contact"] for r in rows),
            "speedup_repeated_median": statistics.median(r["speedup"] for r in rows),
            "speedup_repeated_min": min(r["speedup"] for r in rows),
            "ms_first_contact_learned_median": statistics.median(r["first_contact_ms"]["learned_mg_fcg"][1]
                                                                 for r in rows),
            "ms_first_contact_jacobi_cg1_median": statistics.median(r["first_contact_ms"]["jacobi_cg1_fp64"][1]
                                                                    for r in rows),
            "iterations_learned": sorted({r["learned_mg_fcg"][0] for r in rows}),
            "iterations_jacobi_cg_median": statistics.median(r["jacobi_cg_fp64"][0] for r in rows),
            "ms_learned_median": statistics.median(r["learned_mg_fcg"][1] for r in rows),
            "ms_jacobi_cg_median": statistics.median(r["jacobi_cg_fp64"][1] for r in rows),
            "ms_jacobi_cg1_median": statistics.median(r["jacobi_cg1_fp64"][1] for r in rows),
            "max_relative_residual_learned": max(max(r["relative_residual"][3], r["relative_residual"][4])
                                                 for r in rows),
            "fallback_iterations_total": sum(r["learned_mg_fcg"][3] for r in rows),
            "cpu_baseline": cpu_out}


def cpu_time_to_solve(pattern, vals, rhs, tol=1e-10, reps=3):
    """The reference's CPU solve paths on the same operator and load (copied
    to the host): SciPy CSR Jacobi-CG as ``_scipy_csr_cg`` does it
    (``src/jaxiga/solvers/linear.py:261-321``, restated by the oracle) and
    SuperLU (``spsolve``, the reference's CPU direct default); warm-up, then
    median of ``reps``, one core (SciPy's CSR product and SuperLU are
    single-threaded)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from threadpoolctl import threadpool_limits

    from oracle import solver as OSV

    n = pattern.shape[0]
    A = sp.csr_matrix((vals.detach().cpu().numpy(), pattern.cols_d.cpu().numpy(), pattern.indptr_d.cpu().numpy()),
                      shape=(n, n))
    b = rhs.detach().cpu().numpy()
    bn = np.linalg.norm(b)
    out = {}
    with threadpool_limits(1):
        for name, fn in (("scipy_jacobi_cg", lambda: OSV.scipy_csr_cg(A, b, tol)),
                         ("superlu", lambda: (spla.spsolve(A.tocsc(), b, permc_spec="MMD_AT_PLUS_A"), None))):
            fn()
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                u, it = fn()
                times.append(time.perf_counter() - t0)
            out[name] = {"seconds": statistics.median(times), "iterations": it,
                         "relative_residual": float(np.linalg.norm(b - A @ u) / bn)}
    out.update({"cores": 1, "kind": "port", "host_cpus": os.cpu_count(),
                "sample": f"the full operator of the first unseen seed ({n} free dofs, {pattern.nnz} "
                          f"nonzeros), copied to the host; warm-up, then median of {reps}"})
    return out


def matfree_bench(refine=6, reps=5, log=print):
    """cfg3: 3-D Poisson on the unit cube, p=2, 64^3 elements, a = exp(GRF l=0.1,
    sigma=1) at the Gauss points, source 1, zero Dirichlet. Matrix-free
    sum-factorised apply (FP64, FP32, and BF16 coefficient with FP32
    arithmetic) against the assembled FP64 CSR SpMV, and the solves to 1e-10:
    FP64 Jacobi-CG on CSR, FP64 matrix-free Jacobi-CG, FP32 and BF16
    matrix-free inner solves with FP64 refinement (Jacobi from the
    sum-factorised diagonal)."""
    import torch

    import jaxiga_amd as jx
    from jaxiga_amd import _lib
    from jaxiga_amd import grf
    from jaxiga_amd import linear as LIN
    from jaxiga_amd import matfree as MF
    from jaxiga_amd import methods as M
    from jaxiga_amd import precision as PR

    V = jx.FunctionSpace(jx.primitives.cuboid([0, 0, 0], [1, 1, 1]).elevate(2).refine(refine))
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))], source=1.0)
    ctx = M.build_context(prob)
    a = grf.gauss_point_field(ctx, seed=3, ell=0.1, sigma=1.0)
    a.exp_()
    params = {"a0": jx.PointField(a)}
    vals, b, diag = M.assemble(ctx, params, diagonal=True)
    pattern = ctx.assembly
    dinv = MF.inverse_guard(diag)
    ops = {"fp64": MF.MatrixFreeOperator(ctx, params, _lib.IGA_FP64),
           "fp32": MF.MatrixFreeOperator(ctx, params, _lib.IGA_FP32),
           "bf16": MF.MatrixFreeOperator(ctx, params, _lib.IGA_BF16)}
    x = torch.randn(ctx.n_free, dtype=torch.float64, device="cuda")

    def timed(fn):
        return graph_ms(fn) * 1e-3

    t_csr = timed(lambda: LIN.spmv(pattern, vals, x))
    y_csr = LIN.spmv(pattern, vals, x)
    t_apply = {k: timed(lambda op=op: op.apply(x)) for k, op in ops.items()}
    err = {k: float((op.apply(x) - y_csr).abs().max() / y_csr.abs().max()) for k, op in ops.items()}
    d_sf = ops["fp64"].diagonal()
    diag_err = float((d_sf - diag).abs().max() / diag.abs().max())
    t_diag = timed(lambda: ops["fp64"].diagonal())
    n_e, n_q, n = V.n_elems, 27, ctx.n_free
    # algorithmic bytes of one apply: the coefficient at every Gauss point in
    # its storage type, x read and y written once (FP64)
    coef_bytes = {"fp64": 8, "fp32": 4, "bf16": 2}
    mf_bytes = {k: n_e * n_q * c + 2 * n * 8 for k, c in coef_bytes.items()}
    csr_bytes = pattern.nnz * 12 + n * 8 * 2 + (n + 1) * 4

    def solve_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    dinv_sf = MF.inverse_guard(d_sf)
    jac, t_j = solve_time(lambda: LIN.cg_csr(pattern, vals, b, 1e-10, dinv=dinv))
    mfr, t_m = solve_time(lambda: ops["fp64"].cg(b, 1e-10, dinv=dinv_sf))
    ref32, t_r32 = solve_time(lambda: PR.mf_mixed_solve(ops["fp64"], ops["fp32"], b, dinv_sf, 1e-10))
    ref16, t_r16 = solve_time(lambda: PR.mf_mixed_solve(ops["fp64"], ops["bf16"], b, dinv_sf, 1e-10,
                                                          inner_tol=1e-3))
    bn = LIN.residual_norms(pattern, vals, b * 0, b)[1]
    rel = {k: LIN.residual_norms(pattern, vals, r.x, b)[0] / bn for k, r in
           (("fp64", mfr), ("fp32", ref32), ("bf16", ref16))}
    traffic = measured_traffic(os.path.join(ROOT, "profiles", "matfree_cube_64.md"))
    roof = {f"matfree_{k}": hbm_roofline(
        f"k_mfs_apply + k_mfs_gather <3,2,{k}>", mf_bytes[k], t * 1e3,
        f"{n_e} elements x {n_q} points x {coef_bytes[k]} B coefficient + x read and y written once "
        f"({n} x 8 B each)", traffic.get(f"k_mfs_apply+k_mfs_gather<3,2,{k}>")) for k, t in t_apply.items()}
    roof["csr_fp64"] = hbm_roofline(
        "k_spmv<16,f64>", csr_bytes, t_csr * 1e3,
        f"{pattern.nnz} nonzeros x (8 B value + 4 B column) + row pointers + x read and y written once",
        measured_traffic(os.path.join(ROOT, "profiles", "spmv_csr.md")).get("k_spmv<16,f64,cube-64>"))
    out = {
        "workload": f"unit-cube Poisson, p=2, {2**refine}^3 elements, a=exp(GRF l=0.1 sigma=1), {n} free dofs, "
                    f"{pattern.nnz} nonzeros",
        "roofline": roof["matfree_fp32"],
        "roofline_apply": roof,
        "apply_seconds": {"csr_fp64": t_csr, **{f"matfree_{k}": t for k, t in t_apply.items()}},
        "apply_algorithmic_GBs": {"csr_fp64": csr_bytes / t_csr / 1e9,
                                  **{f"matfree_{k}": mf_bytes[k] / t / 1e9 for k, t in t_apply.items()}},
        "apply_algorithmic_bytes": {"csr_fp64": csr_bytes, **{f"matfree_{k}": b_ for k, b_ in mf_bytes.items()}},
        "apply_rel_error_vs_csr_fp64": {f"matfree_{k}": e for k, e in err.items()},
        "sum_factorised_diagonal": {"seconds": t_diag, "rel_error_vs_assembled": diag_err},
        "solve": {"csr_fp64_jacobi_cg": {"seconds": t_j, "iterations": jac.iterations},
                  "matfree_fp64_jacobi_cg": {"seconds": t_m, "iterations": mfr.iterations,
                                             "relative_residual": rel["fp64"]},
                  "matfree_fp32_refined": {"seconds": t_r32, "outer": ref32.outer, "inner": ref32.total_inner,
                                           "relative_residual": rel["fp32"]},
                  "matfree_bf16_refined": {"seconds": t_r16, "outer": ref16.outer, "inner": ref16.total_inner,
                                           "relative_residual": rel["bf16"]}},
    }
    log(f"[bench] cfg3 apply: csr64 {t_csr * 1e3:.3f} ms, " +
        ", ".join(f"mf {k} {t * 1e3:.3f} ms" for k, t in t_apply.items()) +
        f"; solve csr {t_j * 1e3:.1f} ms ({jac.iterations} it), mf64 {t_m * 1e3:.1f} ms ({mfr.iterations} it), "
        f"mf32+IR {t_r32 * 1e3:.1f} ms, bf16+IR {t_r16 * 1e3:.1f} ms "
        f"(rel.res {rel['fp32']:.1e}, {rel['bf16']:.1e})")
    return out


def batched_bench(n_instances=512, cpu=True, log=print):
    """cfg4: 512 uneven unit-square Poisson instances (p in {2,3}, 16..128
    elements per side, a = exp(GRF), source 1): batched assembly plus the
    batched Jacobi-CG, method and tiles from the autotuner (the segmented
    solve eagerly, the persistent one replayed from a captured graph);
    instances per second, the worst relative residual, the solve's CSR
    traffic against HBM, and the oracle's per-instance time on the CPU."""
    import torch

    from jaxiga_amd import batched as BT

    bp = BT.BatchedPoisson(BT.cfg4_specs(n_instances))
    best, timings = bp.autotune()
    if bp.method == "persistent":
        graph = bp.capture(tol=1e-10)
        job = graph.replay
    else:
        def job():
            bp.run(tol=1e-10)
    job()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    job()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the solve alone (assembled operators resident)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    bp.solve(1e-10)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t1
    iters, rel = bp.results()
    # algorithmic bytes of the solve: per instance and iteration the CSR
    # (8 B value + 4 B column per nonzero, row pointers) and the 13 vector
    # streams of one Jacobi-PCG iteration (FP64)
    n_rows = bp.n_rows.astype(np.int64)
    per_it = bp.nnz * 12 + (n_rows + 1) * 4 + 13 * n_rows * 8
    solve_bytes = int((iters.astype(np.int64) * per_it).sum())
    roof = hbm_roofline(f"batched Jacobi-CG ({bp.method})", solve_bytes, t_solve * 1e3,
                        "sum over instances of iterations x (12 B per nonzero + row pointers + 13 x 8 B "
                        "per row)")
    log(f"[bench] cfg4: {n_instances} instances in {dt * 1e3:.1f} ms ({n_instances / dt:.0f} instances/s), "
        f"solve {t_solve * 1e3:.1f} ms ({roof['achieved']:.0f} GB/s), method {best}, "
        f"max rel.res {float(rel.max()):.2e}, iterations {int(iters.min())}..{int(iters.max())}")
    out = {"workload": f"{n_instances} unit-square Poisson instances, p in {{2,3}}, 16..128 elements per side",
           "seconds": dt, "instances_per_second": n_instances / dt, "method_lanes_per_row_threads": list(best),
           "solve_seconds": t_solve, "roofline": roof,
           "max_relative_residual": float(rel.max()), "iterations_min_max": [int(iters.min()), int(iters.max())],
           "elements": bp.n_elements}
    if cpu:
        out["cpu_baseline"] = batched_cpu_baseline(bp, n_instances)
    return out


def batched_cpu_baseline(bp, n_instances):
    """The oracle's FP64 assembly (basis, element matrices, CSR, load) and
    Jacobi-PCG to 1e-10 of one instance of each (degree, mesh) class of the
    batch, one core, with the batch's coefficient values as input; the batch
    time is the class times weighted by the class sizes."""
    from threadpoolctl import threadpool_limits

    from oracle import assembly as OA
    from oracle import forms as OF
    from oracle import geometry as OG
    from oracle import solver as OSV
    from oracle import space as OS

    everywhere = lambda x: np.ones(x.shape[0], bool)  # noqa: E731
    total, per_class = 0.0, {}
    with threadpool_limits(1):
        for g in bp.groups:
            p, n_el = g.degree, g.n_el
            coef = g.coef[0].cpu().numpy()
            t0 = time.perf_counter()
            O = OG.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(p).refine(int(round(np.log2(n_el))))
            VO = OS.function_space(O)
            presc, vals = OF.resolve_dirichlet(VO, [OF.DirichletBC(0.0, everywhere)])
            dm = OS.make_dofmap(VO.n_dofs, presc, vals)
            basis = OS.gauss_basis(VO)
            params = {"a0": OF.PointField(coef)}
            Ke = OF.poisson_tangents(basis, params)
            pat = OA.csr_pattern(VO.elem_dofs, VO.n_dofs, 1, dm.free)
            data = OA.assemble_csr(pat, Ke)
            gd = OF.gdofs(VO)
            load = OA.global_load(VO.n_dofs, gd, OF.source_load(basis, lambda x, q: 1.0, params, 1))
            b = (load - OA.lifting(VO.n_dofs, gd, Ke, dm))[dm.free]
            K = OA.to_scipy(pat, data)
            OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-10)
            t = time.perf_counter() - t0
            per_class[f"p{p}-n{n_el}"] = t
            total += t * len(g.members)
    return {"value": n_instances / total, "unit": "instances/s", "cores": 1, "kind": "port",
            "sample": "oracle FP64 assembly + Jacobi-PCG to 1e-10 of one instance per (degree, mesh) class "
                      "(" + ", ".join(f"{k} {v:.2f} s" for k, v in per_class.items()) + "), weighted by the "
                      "class sizes of the batch"}


def adjoint_bench(refine=8, reps=3, log=print):
    """Reverse pass of the Galerkin solve (implicit differentiation) on the
    cfg5 Poisson problem: unit square, p=3, 256^2 elements, a = exp(GRF l=0.1
    sigma=1.5) as values at the 16 Gauss points of every element that carry a
    gradient, loss c . u. Forward (assembly + Jacobi-CG to 1e-10) and reverse
    pass (one adjoint Jacobi-CG, basis gradients at the Gauss points, the
    fields-at-points kernel, the point cotangents), medians of ``reps`` after a
    warm-up; the full-size gradient checked by a central difference in a
    random direction; ``k_field_points`` timed alone with HIP events."""
    import torch

    import jaxiga_amd as jx
    from jaxiga_amd import adjoint as AD
    from jaxiga_amd import grf
    from jaxiga_amd import linear as LIN
    from jaxiga_amd.evaluation import evaluate
    from jaxiga_amd.pointset import gauss

    ctx, _, _ = grf.poisson_instance(refine, 0)
    prob, V = ctx.problem, ctx.space
    a = grf.gauss_point_field(ctx, 0, 0.1, 1.5).exp_()
    lin = jx.LinearOptions(method="cg", preconditioner="jacobi", tol=1e-10)
    gen = torch.Generator(device="cuda").manual_seed(7)
    c = torch.randn(V.n_dofs, dtype=torch.float64, device="cuda", generator=gen)

    def run():
        a_t = a.clone().requires_grad_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = jx.solve(prob, params={"a0": jx.PointField(a_t)}, linear=lin)
        loss = (c * sol.u).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        LIN.reset_solve_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return a_t.grad, t1 - t0, t2 - t1, sol, dict(LIN.SOLVE_COUNTER)

    run()
    runs = [run() for _ in range(reps)]
    g, _, _, sol, counter = runs[-1]
    t_fwd = statistics.median(r[1] for r in runs)
    t_bwd = statistics.median(r[2] for r in runs)
    # the same gradient through LinearOptions(preconditioner="multigrid"):
    # multigrid flexible CG (damped-Jacobi smoothers) forward and adjoint
    lin_jacobi, lin = lin, jx.LinearOptions(method="cg", preconditioner="multigrid", tol=1e-10)
    run()
    runs_mg = [run() for _ in range(reps)]
    g_mg = runs_mg[-1][0]
    mg = {"forward_seconds": statistics.median(r[1] for r in runs_mg),
          "reverse_seconds": statistics.median(r[2] for r in runs_mg),
          "solves_per_gradient": runs_mg[-1][4],
          "max_rel_diff_vs_jacobi_gradient": float((g_mg - g).abs().max() / g.abs().max())}
    lin = lin_jacobi
    # the parts of the reverse pass, timed alone
    free = ctx.free_d.long()
    vals, b, _ = jx.methods.assemble(ctx, {"a0": jx.PointField(a)})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam_free = LIN.adjoint_solve(lin, ctx.assembly, vals, c[free])
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    lam = torch.zeros_like(sol.u)
    lam[free] = lam_free
    t0 = time.perf_counter()
    basis = evaluate(V, gauss(V, ctx.n_q1), outputs=("dR", "x", "w"))
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    t_k = event_ms(lambda: AD.field_points(basis, sol.u, lam, 1, values=False)) * 1e-3
    n_e, n_q, dim, n_l = basis.dR.shape
    n_p = n_e * n_q
    k_bytes = n_p * dim * n_l * 8 + n_e * n_l * 4 + 2 * V.n_dofs * 8 + 2 * n_p * dim * 8
    # directional check of the full-size gradient
    d = torch.randn(a.shape, dtype=torch.float64, device="cuda", generator=gen) * a
    h = 1e-4

    def loss_at(aa):
        with torch.no_grad():
            return float((c * jx.solve(prob, params={"a0": jx.PointField(aa)}, linear=lin).u).sum())

    fd = (loss_at(a + h * d) - loss_at(a - h * d)) / (2 * h)
    ad = float((g * d).sum())
    out = {
        "workload": f"unit-square Poisson, p=3, {2**refine}x{2**refine} elements, a=exp(GRF l=0.1 sigma=1.5) "
                    f"at {n_p} Gauss points (the differentiated input), loss c.u, Jacobi-CG to 1e-10",
        "elements": n_e, "forward_seconds": t_fwd, "reverse_seconds": t_bwd,
        "reverse_parts_seconds": {"adjoint_cg": t_adj, "basis_gradients": t_basis, "field_points": t_k},
        "solves_per_gradient": counter,
        "multigrid_preconditioner": mg,
        "directional_check": {"adjoint": ad, "central_difference": fd, "rel_diff": abs(ad - fd) / abs(fd)},
        "roofline": {"bound": "hbm", "kernel": "k_field_points_lds<1,2,2>", "kernel_ms": t_k * 1e3,
                     "achieved": k_bytes / t_k / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": k_bytes / t_k / 1e9 / 8000.0,
                     "algorithmic": f"{n_p} points x ({dim}x{n_l} basis gradients x 8 B + 2 fields x {dim} "
                                    f"gradients x 8 B) + {n_e}x{n_l} dof indices x 4 B + 2 x {V.n_dofs} "
                                    f"coefficients x 8 B",
                     "traffic": measured_traffic(os.path.join(ROOT, "profiles", "adjoint_poisson_256.md")).get(
                         "k_field_points_lds<1,2,2>")},
    }
    log(f"[bench] adjoint: forward {t_fwd * 1e3:.1f} ms, reverse {t_bwd * 1e3:.1f} ms (adjoint CG "
        f"{t_adj * 1e3:.1f} ms, basis {t_basis * 1e3:.2f} ms, k_field_points {t_k * 1e6:.1f} us = "
        f"{k_bytes / t_k / 1e9:.0f} GB/s); solves {counter}; d/dir {ad:.8e} vs central difference {fd:.8e}; "
        f"with preconditioner='multigrid': forward {mg['forward_seconds'] * 1e3:.1f} ms, reverse "
        f"{mg['reverse_seconds'] * 1e3:.1f} ms, gradient {mg['max_rel_diff_vs_jacobi_gradient']:.1e} from Jacobi's")
    return out


def adjoint_cpu_baseline(sample_refine=6):
    """The oracle's reverse pass (adjoint direct solve, element and point
    cotangents) for the same Poisson problem at 2**sample_refine elements per
    direction, one core; median of 3 after a warm-up."""
    from threadpoolctl import threadpool_limits

    from oracle import adjoint as OJ
    from oracle import forms as OF
    from oracle import geometry as OG
    from oracle import grf as OGR
    from oracle import pipeline as OP
    from oracle import space as OS

    VO = OS.function_space(OG.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(3).refine(sample_refine))
    x = OS.gauss_basis(VO).x
    om, ph = OGR.rff_params(0, 2, 0.1)
    a0 = np.exp(OGR.rff_field(x.reshape(-1, 2), om, ph, 1.5).reshape(x.shape[:2]))
    out = OP.solve("poisson", VO, [OF.DirichletBC(0.0, lambda xx: np.ones(len(xx), bool))],
                   source=lambda xx, p: 1.0 + 0.0 * xx[0], params={"a0": OF.PointField(a0)}, method="direct")
    c = np.random.default_rng(7).standard_normal(VO.n_dofs)
    gd = OF.gdofs(VO)
    dm = out["dofmap"]

    def once():
        db, lam = OJ.csr_solve_vjp(out["pattern"], out["data"], out["u_free"], c[dm.free])
        Kbar, lb = OJ.element_cotangent(out["pattern"], db, lam, gd, dm)
        return OJ.point_cotangents("poisson", out["basis"], Kbar, lb, gd, 1)

    with threadpool_limits(1):
        once()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            once()
            times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {"value": VO.n_elems / t / 1e6, "unit": "Melem/s (reverse pass)", "cores": 1, "kind": "port",
            "sample": f"oracle reverse pass (adjoint direct solve + element and point cotangents) of the "
                      f"Poisson problem at {2**sample_refine}x{2**sample_refine} elements, median of 3 runs "
                      f"of {t:.3f} s"}


def assembly_run(ctx, params, prec_name, steps, warmup):
    """Time ``steps`` full assemblies of the operator (element pass, row gather,
    load with its boundary basis) with inputs resident; HIP events on the
    stream the kernels run on give the per-kernel durations."""
    import torch

    from jaxiga_amd import _lib
    from jaxiga_amd import methods as M

    precision = _lib.IGA_FP32 if prec_name == "fp32" else _lib.IGA_FP64
    dtype = "f32" if prec_name == "fp32" else "f64"
    peak_tf = PEAK_FP32_TFLOPS if prec_name == "fp32" else PEAK_FP64_TFLOPS
    pattern = ctx.assembly
    n_e = ctx.space.n_elems
    # arguments and output buffers are prepared once; every step recomputes
    # everything (basis, element matrices, gather, load with its boundary
    # basis), all on this stream; the timing events are created before the
    # timed region.
    job = M.prepare_assembly(ctx, params, precision=precision)

    def step():
        job.rhs.zero_()
        M.run_assembly(ctx, params, job, passes=(1, 2))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # per-kernel durations: each pass launched back to back on this stream
    # between HIP events (median over batches)
    t_el = event_ms(lambda: job.launch(ctx, passes=1))
    t_ga = event_ms(lambda: job.launch(ctx, passes=2))
    # the operator and load of one full step, for the checks below
    step()
    torch.cuda.synchronize()
    vbytes = 4 if precision == _lib.IGA_FP32 else 8
    flops = elasticity_flops_per_element() * n_e
    el = {"bound": "mfma", "achieved": flops / (t_el * 1e-3) / 1e12, "peak": peak_tf, "unit": "TFLOP/s",
          "kernel": f"k_element_w<2,3,elasticity,{dtype}>", "kernel_ms": t_el,
          "algorithmic": f"{elasticity_flops_per_element()} flop/element x {n_e} elements"}
    el["frac"] = el["achieved"] / peak_tf
    # FP64 element matrices are stored as their upper 2x2 blocks (136 of 256)
    ke_vals, ke_desc = (136 * 4, "136 upper 2x2 blocks") if precision == _lib.IGA_FP64 else (32 * 32, "32x32")
    # the gather finds each value's CSR slot by a ballot over the stencil
    # box: it reads no column indices, only the row pointers
    gbytes = n_e * ke_vals * vbytes + pattern.nnz * vbytes + (ctx.n_free + 1) * 4
    ga = {"bound": "hbm", "achieved": gbytes / (t_ga * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
          "kernel": f"k_gather_points<2,3,elasticity,{dtype}>", "kernel_ms": t_ga,
          "algorithmic": f"element matrices {n_e}x{ke_desc}x{vbytes} B read + {pattern.nnz} values x "
                         f"{vbytes} B written + row pointers"}
    ga["frac"] = ga["achieved"] / PEAK_HBM_GBS
    traffic = measured_traffic()
    el["traffic"] = traffic.get(el["kernel"])
    ga["traffic"] = traffic.get(ga["kernel"])
    dominant, secondary = (el, ga) if t_el >= t_ga else (ga, el)
    value = n_e * steps / elapsed / 1e6
    # whole assembly against max(compute, compulsory bytes): control points
    # and weights read, CSR values and load written once
    n_cp = ctx.space.n_dofs // ctx.space.vec
    comp_bytes = n_cp * 3 * 8 + pattern.nnz * vbytes + ctx.n_free * 8
    roof_ms = max(flops / (peak_tf * 1e12), comp_bytes / (PEAK_HBM_GBS * 1e9)) * 1e3
    overall = {"bound": "mfma" if flops / (peak_tf * 1e12) * 1e3 >= roof_ms else "hbm",
               "roof_ms": roof_ms, "ms_per_step": elapsed / steps * 1e3,
               "frac": roof_ms / (elapsed / steps * 1e3),
               "algorithmic": f"{flops} flop, {comp_bytes} compulsory bytes per assembly"}
    print(f"[bench] {dtype} assembly {elapsed / steps * 1e3:.3f} ms/step (element {t_el:.3f} ms, "
          f"gather {t_ga:.3f} ms), {value:.1f} Melem/s", flush=True)
    return {"value": value, "ms_per_step": elapsed / steps * 1e3, "dtype": dtype, "roofline": dominant,
            "roofline_secondary": secondary, "roofline_assembly": overall, "vals": job.vals,
            "rhs": job.rhs}


def _lib_mod():
    from jaxiga_amd import _lib

    return _lib.load()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    # the GPU needs a few milliseconds of work to reach its steady clock: the
    # first steps of a cold run are up to 12 % slower
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-solve measurement")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32",
                    help="assembly precision of the headline value (the other is reported beside it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (HBM counter calibration)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner solve")
    ap.add_argument("--learned-only", action="store_true",
                    help="only run the learned-preconditioner time-to-solve comparison")
    ap.add_argument("--learned-fields", type=int, default=88,
                    help="training fields of the learned PC (about 23 residual/error pairs each)")
    ap.add_argument("--learned-steps", type=int, default=300, help="training steps of the learned PC")
    ap.add_argument("--learned-seeds", type=int, default=32,
                    help="unseen coefficient seeds (10000, 10001, ...) of the learned-PC comparison")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the matrix-free (3-D) and batched sections")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the reverse-pass (implicit differentiation) measurement")
    ap.add_argument("--extras-only", action="store_true",
                    help="only run the matrix-free (3-D) and batched sections")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import jaxiga_amd as jx
    from jaxiga_amd import _lib
    from jaxiga_amd import linear as LIN
    from jaxiga_amd import methods as M

    torch.cuda.init()
    if args.copy_calibration:
        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] copied 5 x 256 MiB", flush=True)
        return
    lib = _lib.load()
    if args.learned_only:
        print(json.dumps({"learned_pc": learned_pc_bench(
            n_train=args.learned_fields, steps=args.learned_steps,
            test_seeds=tuple(range(10_000, 10_000 + args.learned_seeds)), cpu=not args.no_cpu)}), flush=True)
        return
    if args.adjoint_only:
        adj = adjoint_bench()
        if not args.no_cpu:
            adj["cpu_baseline"] = adjoint_cpu_baseline()
            adj["gpu_reverse_Melem_per_s"] = adj["elements"] / adj["reverse_seconds"] / 1e6
        print(json.dumps({"adjoint": adj}), flush=True)
        return
    if args.extras_only:
        print(json.dumps({"matfree": matfree_bench(), "batched": batched_bench(cpu=not args.no_cpu)}), flush=True)
        return
    V, prob, params = build_problem()
    ctx = M.build_context(prob, params)
    pattern = ctx.assembly
    n_e = V.n_elems
    print(f"[bench] elements={n_e} dofs={V.n_dofs} free={ctx.n_free} nnz={pattern.nnz}", flush=True)
    other = "fp64" if args.precision == "fp32" else "fp32"
    main_run = assembly_run(ctx, params, args.precision, args.steps, args.warmup)
    side_run = assembly_run(ctx, params, other, args.steps, args.warmup)
    runs = {args.precision: main_run, other: side_run}
    value, ms_per_step, dtype = main_run["value"], main_run["ms_per_step"], main_run["dtype"]
    dominant, secondary = main_run["roofline"], main_run["roofline_secondary"]
    v64, v32 = runs["fp64"]["vals"], runs["fp32"]["vals"]
    stiff_err = float((v32.double() - v64).abs().max() / v64.abs().max())
    print(f"[bench] FP32 vs FP64 stiffness entries: max |dK| / max |K| = {stiff_err:.2e}", flush=True)

    # -- time to solve (rel. residual 1e-10): FP64 Jacobi-CG, and FP32 operator
    # -- with FP64 iterative refinement (outer residual on the FP64 operator)
    solve = spmv_roof = None
    if not args.no_solve:
        from jaxiga_amd import precision as PR

        rhs = runs["fp64"]["rhs"]
        dinv = LIN.inverse_diagonal(pattern, v64)
        work = torch.empty(int(lib.iga_cg_workspace(ctx.n_free)), dtype=torch.float64, device="cuda")
        LIN.cg_csr(pattern, v64, rhs, 1e-10, maxiter=50, dinv=dinv, workspace=work)  # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = LIN.cg_csr(pattern, v64, rhs, 1e-10, maxiter=20 * ctx.n_free, dinv=dinv, workspace=work)
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t1
        rn, bn = LIN.residual_norms(pattern, v64, res.x, rhs)
        # the CSR SpMV of the solve against HBM: FP64 values, column indices
        # and row pointers read once, x read and y written once
        xs = torch.randn(ctx.n_free, dtype=torch.float64, device="cuda")
        t_sp = graph_ms(lambda: LIN.spmv(pattern, v64, xs))
        sp_bytes = pattern.nnz * 12 + (ctx.n_free + 1) * 4 + 2 * ctx.n_free * 8
        spmv_roof = hbm_roofline(
            "k_spmv<16,f64>", sp_bytes, t_sp,
            f"{pattern.nnz} nonzeros x (8 B value + 4 B column) + {ctx.n_free + 1} row pointers x 4 B "
            f"+ x read and y written once ({ctx.n_free} x 8 B each)",
            measured_traffic(os.path.join(ROOT, "profiles", "spmv_csr.md")).get("k_spmv<16,f64,elasticity-256>"))
        # the matrix-free linear-elasticity operator of the same problem
        # (rational patch: per Gauss point sqrt(w|det J|) J^-1 and W, grad W)
        from jaxiga_amd import matfree as MF

        y_csr = LIN.spmv(pattern, v64, xs)
        mf_el = {"csr_spmv_ms": t_sp}
        ops = {}
        for name, prec, tb in (("fp64", _lib.IGA_FP64, 8), ("fp32", _lib.IGA_FP32, 4)):
            op = ops[name] = MF.MatrixFreeOperator(ctx, params, prec)
            t_ap = graph_ms(lambda op=op: op.apply(xs))
            err = float((op.apply(xs) - y_csr).abs().max() / y_csr.abs().max())
            ap_bytes = n_e * 16 * (4 + 3) * tb + 2 * ctx.n_free * 8
            mf_el[name] = {"apply_ms": t_ap, "rel_err_vs_csr_fp64": err, "roofline": hbm_roofline(
                f"k_mf_apply_el + k_mf_gather <2,3,{name}>", ap_bytes, t_ap,
                f"{n_e} elements x 16 points x 7 factors x {tb} B (J^-1 scaled, W, grad W) + x read "
                f"and y written once")}
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res_mf = ops["fp64"].cg(rhs, 1e-10, maxiter=20 * ctx.n_free, dinv=dinv)
        torch.cuda.synchronize()
        mf_el["solve_fp64_jacobi_cg"] = {"seconds": time.perf_counter() - t1, "iterations": res_mf.iterations}
        del ops
        # FP32 operator, FP64 vectors, residual replaced with the FP64 one
        PR.csr_mixed_cg(pattern, v64, v32, rhs, dinv, tol=1e-10, maxiter=50, workspace=work)
        torch.cuda.synchronize()
        mixed = PR.csr_mixed_cg(pattern, v64, v32, rhs, dinv, tol=1e-10,
                                maxiter=20 * ctx.n_free, workspace=work)
        rn_m, _ = LIN.residual_norms(pattern, v64, mixed.x, rhs)
        disp_err = float((mixed.x - res.x).abs().max() / res.x.abs().max())
        # FP32 inner solves with FP64 iterative refinement, for comparison
        refined = PR.csr_mixed_solve(pattern, v64, v32, rhs, dinv, tol=1e-10)
        rn_r, _ = LIN.residual_norms(pattern, v64, refined.x, rhs)
        solve = {"fp64_jacobi_cg": {"seconds": t_solve, "iterations": res.iterations,
                                    "relative_residual_fp64": rn / bn},
                 "fp32_operator_reliable_updates": {
                     "seconds": mixed.seconds, "iterations": mixed.iterations,
                     "relative_residual_fp64": rn_m / bn,
                     "max_rel_displacement_diff_vs_fp64": disp_err},
                 "fp32_operator_fp64_refinement": {"seconds": refined.seconds, "outer": refined.outer,
                                                   "inner_iterations": refined.total_inner,
                                                   "relative_residual_fp64": rn_r / bn},
                 "fp32_vs_fp64_stiffness_max_rel": stiff_err,
                 "matrix_free_elasticity": mf_el}
        print(f"[bench] jacobi-cg fp64 {t_solve:.3f} s, {res.iterations} it, rel.res {rn / bn:.2e}; "
              f"fp32 + reliable updates {mixed.seconds:.3f} s, {mixed.iterations} it, "
              f"rel.res {rn_m / bn:.2e}, displacement diff {disp_err:.2e}; fp32 + refinement "
              f"{refined.seconds:.3f} s ({refined.outer} outer, {refined.total_inner} inner); matrix-free "
              f"apply fp64 {mf_el['fp64']['apply_ms']:.4f} ms, fp32 {mf_el['fp32']['apply_ms']:.4f} ms "
              f"(CSR {t_sp:.4f} ms), matrix-free CG {mf_el['solve_fp64_jacobi_cg']['seconds']:.3f} s "
              f"({mf_el['solve_fp64_jacobi_cg']['iterations']} it)", flush=True)

    learned = None if (args.no_learned or args.no_solve) else learned_pc_bench(
        n_train=args.learned_fields, steps=args.learned_steps,
        test_seeds=tuple(range(10_000, 10_000 + args.learned_seeds)), cpu=not args.no_cpu)
    extras = {} if (args.no_extras or args.no_solve) else {"matfree": matfree_bench(),
                                                           "batched": batched_bench(cpu=not args.no_cpu)}
    if not (args.no_extras or args.no_solve):
        adj = adjoint_bench()
        if not args.no_cpu:
            adj["cpu_baseline"] = adjoint_cpu_baseline()
            adj["gpu_reverse_Melem_per_s"] = adj["elements"] / adj["reverse_seconds"] / 1e6
        extras["adjoint"] = adj
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "assembly Melem/s; time-to-solve @1e-10 rel res, learned-PC vs plain CG",
        "value": value,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": "quarter-annulus plane-strain elasticity, p=3, 256x256 elements, "
                               f"134162 dofs, {dtype} assembly (basis + K_e + CSR scatter + load); "
                               "solve on the FP32 operator with FP64 residual replacement",
                   "elements": n_e, "nnz": pattern.nnz, "n_free": ctx.n_free},
        "roofline": dominant,
        "roofline_secondary": secondary,
        "roofline_assembly": main_run["roofline_assembly"],
        "roofline_spmv": spmv_roof,
        "cpu_baseline": cpu,
        "time_to_solve": solve,
        f"assembly_{other}": {k: side_run[k] for k in ("value", "ms_per_step", "dtype", "roofline",
                                                          "roofline_secondary", "roofline_assembly")},
        "learned_pc": learned,
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
