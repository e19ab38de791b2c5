"""Batched assembly + batched CG, segmented and persistent (the reference
batches by vmap and checks batch == loop, tests/test_advanced.py:337-344).

* Every (degree, mesh) class of the 512-instance workload -- p in {2, 3} x
  {16, 32, 64, 128} elements per side, exp(GRF) coefficients -- against the
  oracle's FP64 assembly and solve: CSR values and load within 1e-12 of
  their scale, Jacobi-PCG iteration counts within 2 of the oracle's, the
  solution within 1e-8 of the oracle's sparse direct solution relative to
  its scale (CG to 1e-12), the oracle's FP64 residual of the GPU solution
  <= 2e-12.
* The full 512-instance batch: every independent FP64 residual <= 1e-10.
* All of the above for both solve methods (segmented: every instance's rows
  spread over the GPU; persistent: one workgroup per instance).
* Every instance of an uneven batch matches the single-instance path.
* The autotuner picks a valid candidate; hipGraph replays reproduce the eager
  result (the load vector is summed with atomics, so replays agree to
  rounding, not bitwise).
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import jaxiga_amd as jx
from jaxiga_amd import batched as BT
from jaxiga_amd import grf
from jaxiga_amd import linear as LIN
from jaxiga_amd import methods as M
from oracle import assembly as OA
from oracle import forms as OF
from oracle import grf as OGRF
from oracle import solver as OSV
from oracle import space as OS
from tests import cases as C

pytestmark = pytest.mark.gpu


def _specs():
    rng = np.random.default_rng(3)
    out = []
    for i in range(10):
        out.append(BT.InstanceSpec(int(rng.choice([2, 3])), int(rng.choice([4, 8, 16])),
                                   float(rng.uniform(0.05, 0.2)), 777 + i))
    return out


def _single(spec, sigma=1.0):
    V = jx.FunctionSpace(jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]])
                         .elevate(spec.degree).refine(int(np.log2(spec.n_el))))
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))],
                      source=1.0)
    ctx = M.build_context(prob)
    om, ph = grf.rff_params(spec.seed, 2, spec.ell)
    a = torch.exp(grf.rff_eval(ctx.gauss_points(), om, ph, sigma))
    vals, rhs, _ = M.assemble(ctx, {"a0": jx.PointField(a)})
    return ctx, vals, rhs


METHODS = ["segmented", "persistent"]


@pytest.mark.parametrize("method", METHODS)
def test_batch_matches_single_instances(gpu, method):
    specs = _specs()
    bp = BT.BatchedPoisson(specs)
    bp.run(tol=1e-12, method=method)
    torch.cuda.synchronize()
    iters, rel = bp.results()
    assert (rel <= 1e-12).all(), rel
    for i, s in enumerate(specs):
        print(f"[batched] instance {i} p={s.degree} n={s.n_el}", flush=True)
        ctx, vals, rhs = _single(s)
        g, k = bp.instance(i)
        assert torch.allclose(g.vals[k], vals, rtol=1e-13, atol=1e-13 * float(vals.abs().max()))
        assert torch.allclose(g.rhs[k], rhs, rtol=1e-13, atol=1e-13 * float(rhs.abs().max()))
        ref = LIN.cg_csr(ctx.assembly, vals, rhs, 1e-12)
        assert abs(ref.iterations - int(iters[i])) <= 1, (ref.iterations, iters[i])
        x = g.x[k]
        err = float((x - ref.x).abs().max() / ref.x.abs().max())
        assert err < 1e-8, (i, err)
        r = rhs - LIN.spmv(ctx.assembly, vals, x)
        assert float(r.norm() / rhs.norm()) <= 2e-12


def test_autotune_and_graph_replay(gpu):
    specs = _specs()
    bp = BT.BatchedPoisson(specs)
    cands = [("segmented", 8, 256), ("segmented", 16, 256), ("persistent", 8, 256), ("persistent", 16, 512),
             ("persistent", 32, 1024)]
    best, timings = bp.autotune(candidates=cands, probe_iters=10)
    assert best in timings and len(timings) == 5
    bp.lpr, bp.threads = 16, 512
    bp.run(tol=1e-10, method="persistent")
    torch.cuda.synchronize()
    eager = [g.x.clone() for g in bp.groups]
    it_eager = bp.iters.clone()
    graph = bp.capture(tol=1e-10)
    for rep in range(3):
        for g in bp.groups:
            g.x.fill_(float("nan"))
        bp.iters.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for g, x0 in zip(bp.groups, eager):
            assert torch.allclose(g.x, x0, rtol=1e-9, atol=1e-9 * float(x0.abs().max())), rep
        assert int((bp.iters - it_eager).abs().max()) <= 1
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays


def _oracle_instance(spec, a_gpu, sigma=1.0, n_check=4000):
    """The oracle's FP64 system of one batch instance. The GRF coefficient is
    evaluated once and shipped to both sides: the GPU's values ``a_gpu`` (n_e,
    n_q) are checked against the oracle's own evaluation of the same seeded
    draws at ``n_check`` random (element, point) pairs of the oracle's Gauss
    points (this pins the values and their point order), then used as the
    oracle's PointField; CSR assembly and load follow."""
    _, O = C.patches("square", spec.degree, int(np.log2(spec.n_el)))
    VO = OS.function_space(O)
    presc, vals = OF.resolve_dirichlet(VO, [OF.DirichletBC(0.0, C.everywhere)])
    dm = OS.make_dofmap(VO.n_dofs, presc, vals)
    basis = OS.gauss_basis(VO)
    om, ph = OGRF.rff_params(spec.seed, 2, spec.ell)
    flat_x = basis.x.reshape(-1, 2)
    pick = np.random.default_rng(spec.seed).choice(len(flat_x), min(n_check, len(flat_x)), replace=False)
    a_ref = np.exp(OGRF.rff_field(flat_x[pick], om, ph, sigma))
    a = np.asarray(a_gpu, float)
    assert a.shape == basis.x.shape[:2]
    assert np.abs(a.reshape(-1)[pick] - a_ref).max() <= 1e-12 * np.abs(a_ref).max()
    params = {"a0": OF.PointField(a)}
    Ke = OF.poisson_tangents(basis, params)
    pat = OA.csr_pattern(VO.elem_dofs, VO.n_dofs, 1, dm.free)
    data = OA.assemble_csr(pat, Ke)
    gd = OF.gdofs(VO)
    load = OA.global_load(VO.n_dofs, gd, OF.source_load(basis, lambda x, p: 1.0, params, 1))
    load -= OA.lifting(VO.n_dofs, gd, Ke, dm)
    return pat, data, load[dm.free]


@pytest.mark.parametrize("method", METHODS)
def test_every_batch_class_matches_oracle(gpu, method):
    rng = np.random.default_rng(11)
    specs = [BT.InstanceSpec(p, n, float(rng.uniform(0.05, 0.2)), 5000 + 10 * p + n)
             for p in (2, 3) for n in (16, 32, 64, 128)]
    bp = BT.BatchedPoisson(specs)
    bp.run(tol=1e-12, method=method)
    torch.cuda.synchronize()
    iters, rel = bp.results()
    assert (rel <= 1e-12).all(), rel
    for i, s in enumerate(specs):
        print(f"[batched-oracle] {method} instance {i} p={s.degree} n={s.n_el} ell={s.ell:.3f}", flush=True)
        g, k = bp.instance(i)
        pat, data, b = _oracle_instance(s, g.coef[k].cpu().numpy())
        np.testing.assert_array_equal(g.pattern.indptr, pat.indptr)
        np.testing.assert_array_equal(g.pattern.column_indices, pat.column_indices)
        vals = g.vals[k].cpu().numpy()
        rhs = g.rhs[k].cpu().numpy()
        assert np.abs(vals - data).max() <= 1e-12 * np.abs(data).max()
        assert np.abs(rhs - b).max() <= 1e-12 * np.abs(b).max()
        K = OA.to_scipy(pat, data)
        _, it_ref = OSV.jacobi_pcg(lambda v: K @ v, b, K.diagonal(), tol=1e-12)
        assert abs(it_ref - int(iters[i])) <= 2, (it_ref, int(iters[i]))
        x = g.x[k].cpu().numpy()
        x_ref = spla.spsolve(K.tocsc(), b)
        assert np.abs(x - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
        assert np.linalg.norm(b - K @ x) <= 2e-12 * np.linalg.norm(b)


@pytest.mark.parametrize("method", METHODS)
def test_full_batch_independent_residuals(gpu, method):
    specs = BT.cfg4_specs(512)
    bp = BT.BatchedPoisson(specs)
    bp.lpr, bp.threads = 8, 1024
    print(f"[batched-512] run {method}", flush=True)
    bp.run(tol=1e-10, method=method)
    torch.cuda.synchronize()
    iters, rel = bp.results()
    assert (rel <= 1e-10).all(), float(rel.max())
    worst = 0.0
    for g in bp.groups:
        for k in range(len(g.members)):
            rn, bn = LIN.residual_norms(g.pattern, g.vals[k], g.x[k], g.rhs[k])
            worst = max(worst, rn / bn)
    print(f"[batched-512] worst independent relative residual {worst:.3e}", flush=True)
    assert worst <= 1e-10
