"""Learned preconditioner: multigrid with learned smoothers in flexible CG.

The reference has no learned preconditioner -- ``LinearOptions`` allows only
``"none"`` and ``"jacobi"`` (``src/jaxiga/solvers/linear.py:113-116``) -- so
this is an addition; plain Jacobi-CG stays the fallback and the judge of
convergence (the final FP64 residual is recomputed independently and, if the
flexible solve has not reached the tolerance, Jacobi-CG finishes the job).

Preconditioner: one FP32 V-cycle of the Galerkin multigrid of
:mod:`jaxiga_amd.multigrid`, whose smoother on every level is a stencil of
width ``ks`` per control point, ``S v = D^-1/2 T_W(D^-1/2 v)``, applied once
before and (transposed) once after the coarse correction. The stencil weights
``W`` come from a small convolutional network on the control-point grid -- the
generator -- whose input is the operator's own stencil at every control point
(``K[i, i+o] / K[i, i]``), so the smoother follows the coefficient field; the
generator runs once per level when an operator is set up, not per iteration.

Training (on the GPU, with this package's own forward/backward/Adam kernels):
the solver generates residual/error pairs -- errors of truncated Jacobi-CG
and multigrid-FCG iterates against the converged solution, carried to the
coarser levels by restriction, and errors of damped-Jacobi sweeps from random
starts on every level (the rough components) -- which are coarse-corrected
(``e_c = e - P V(P^T K e)``) and normalised in the energy norm. The loss is the energy of the error left after the
post-smoother, ``mean ||e_c - S^T K e_c||_K^2``, i.e. exactly what the
smoother contributes to the V-cycle's error reduction. The untrained
generator (zero-initialised weights of its last layer) yields damped Jacobi,
which is the method without the network. The generator provides the
smoother of the finest level; the coarser levels keep damped Jacobi (learned
smoothers there, trained on the few points of the small grids, generalised
worse to unseen coefficient fields and raised the iteration count).
"""

from __future__ import annotations

import dataclasses
import math
import time
import weakref

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd import linear as LIN
from jaxiga_amd import multigrid as MG
from jaxiga_amd.runtime import lib, side_stream, stream

KS = 7  # smoother stencil width
OMEGA0 = 0.5  # damped-Jacobi weight of the untrained generator (and of the training sweeps)
OMEGA_COARSE = 0.7  # damped-Jacobi weight of the levels below the generator's, and of the tail
LEARNED_LEVELS = 1  # levels (from the finest) whose smoother the generator provides


def generator_arch(c_in, ks=KS, width=32):
    """(name, kind, c_in, c_out, stride, input, skip, bias, relu); kind 0 = conv."""
    return (
        ("g0", 0, c_in, width, 1, "x", None, True, True),
        ("g1", 0, width, width, 1, "g0", None, True, True),
        ("w", 0, width, ks * ks, 1, "g1", None, True, False),
    )


def _wshape(layer):
    _, kind, cin, cout = layer[:4]
    return (cout, cin, 3, 3) if kind == 0 else (cin, cout, 3, 3)


class ConvNet:
    """A convolutional network: flat FP32 parameters (each layer's weights then
    bias), Adam state and the C descriptor per input size."""

    def __init__(self, arch, seed=0, device="cuda", last_scale=0.0, center_bias=OMEGA0):
        import torch

        self.arch = tuple(arch)
        self.c_in = self.arch[0][2]
        self.offsets, self.bias_offsets = {}, {}
        off = 0
        for layer in self.arch:
            self.offsets[layer[0]] = (off, _wshape(layer))
            off += int(np.prod(_wshape(layer)))
            if layer[7]:
                self.bias_offsets[layer[0]] = off
                off += layer[3]
        self.n = off
        rng = np.random.default_rng(seed)
        w = np.zeros(off, np.float32)
        for i, layer in enumerate(self.arch):
            o, shape = self.offsets[layer[0]]
            fan = layer[2] * 9
            scale = math.sqrt(2.0 / fan) * (last_scale if i == len(self.arch) - 1 else 1.0)
            w[o: o + int(np.prod(shape))] = rng.normal(size=int(np.prod(shape))) * scale
        last = self.arch[-1]
        if last[7] and center_bias:
            w[self.bias_offsets[last[0]] + last[3] // 2] = center_bias
        self.params = torch.as_tensor(w, device=device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._structs = {}

    def weights(self, name):
        o, shape = self.offsets[name]
        return self.params[o: o + int(np.prod(shape))].view(shape)

    def bias(self, name):
        o = self.bias_offsets[name]
        cout = [l for l in self.arch if l[0] == name][0][3]
        return self.params[o: o + cout]

    def load_numpy(self, params: dict):
        import torch

        for name, arr in params.items():
            if name.endswith(".b"):
                self.bias(name[:-2]).copy_(torch.as_tensor(np.asarray(arr, np.float32)))
            else:
                self.weights(name).copy_(torch.as_tensor(np.asarray(arr, np.float32)))

    def struct(self, H, W, precision=0):
        key = (H, W, precision)
        if key in self._structs:
            return self._structs[key]
        s = _lib.IgaNet()
        s.n_layers = len(self.arch)
        s.c_in = self.c_in
        s.precision = precision
        names = [l[0] for l in self.arch]
        sizes = {"x": (H, W)}
        s.H[0], s.W[0] = H, W
        for i, (name, kind, cin, cout, stride, src, skip, bias, relu) in enumerate(self.arch):
            s.kind[i], s.cin[i], s.cout[i], s.stride[i] = kind, cin, cout, stride
            s.src[i] = -1 if src == "x" else names.index(src)
            s.skip[i] = -1 if skip is None else names.index(skip)
            s.woff[i] = self.offsets[name][0]
            s.boff[i] = self.bias_offsets.get(name, -1)
            s.relu[i] = 1 if relu else 0
            h, w = sizes[src]
            if kind == 0:
                oh, ow = (h - 1) // stride + 1, (w - 1) // stride + 1
            else:
                oh, ow = sizes[skip]
            sizes[name] = (oh, ow)
            s.H[i + 1], s.W[i + 1] = oh, ow
        self._structs[key] = s
        return s

    def forward(self, x, precision=0):
        """``x (B, c_in, H, W)`` float32 on the GPU -> ``(output view, activations)``.
        ``precision=1`` runs the stride-1 convolutions with FP16 operands on
        the matrix cores (inference)."""
        import torch

        B, C, H, W = x.shape
        if C != self.c_in:
            raise ValueError(f"network expects {self.c_in} input channels, got {C}")
        st = self.struct(H, W, precision)
        L = lib()
        act = torch.empty(int(L.iga_net_act_floats(st, B)), dtype=torch.float32, device=x.device)
        act[: x.numel()].copy_(x.reshape(-1))
        _lib.check(L.iga_net_forward(st, B, _lib.ptr(self.params), _lib.ptr(act), stream()),
                   "iga_net_forward")
        out_off = int(L.iga_net_act_offset(st, B, len(self.arch) - 1))
        cout = self.arch[-1][3]
        Ho, Wo = st.H[len(self.arch)], st.W[len(self.arch)]
        return act[out_off: out_off + B * cout * Ho * Wo].view(B, cout, Ho, Wo), act

    def weights_version(self):
        """Changes whenever the parameters may have: Adam steps (written
        through the library), in-place torch writes, a new tensor."""
        return (self.step, self.params.data_ptr(), self.params._version)

    def _act(self, H, W, precision, device):
        import torch

        act = self.__dict__.setdefault("_acts", {}).get((H, W, precision))
        if act is None:
            st = self.struct(H, W, precision)
            act = torch.empty(int(lib().iga_net_act_floats(st, 1)), dtype=torch.float32, device=device)
            self._acts[(H, W, precision)] = act
        return act

    def pack16(self, H, W):
        """Pack the FP16 weights that inference at (H, W) reads (into the
        activation buffer kept for that size), on the current stream."""
        st = self.struct(H, W, 1)
        act = self._act(H, W, 1, self.params.device)
        _lib.check(lib().iga_net_pack16(st, 1, _lib.ptr(self.params), _lib.ptr(act), stream()),
                   "iga_net_pack16")
        self.__dict__.setdefault("_packed", {})[(H, W)] = self.weights_version()

    def infer(self, x, out, precision=1):
        """Inference of one sample without copies: input ``x (c_in, H, W)`` and
        output ``out`` (``(c_out, H, W)`` floats) are read and written in place;
        the hidden activations use a buffer kept per input size. With FP16
        operands (``precision=1``) the packed FP16 weights are kept with that
        buffer and packed again only when the parameters have changed."""
        C, H, W = x.shape[-3:]
        if C != self.c_in:
            raise ValueError(f"network expects {self.c_in} input channels, got {C}")
        st = self.struct(H, W, precision)
        act = self._act(H, W, precision, x.device)
        if out.numel() != self.arch[-1][3] * st.H[len(self.arch)] * st.W[len(self.arch)]:
            raise ValueError("output buffer size does not match the network")
        packed = self.__dict__.setdefault("_packed", {})
        if precision == 1 and packed.get((H, W)) == self.weights_version():
            st = self.struct(H, W, 2)  # the same plan, without the packing launches
        _lib.check(lib().iga_net_infer(st, 1, _lib.ptr(self.params), _lib.ptr(act), _lib.ptr(x),
                                       _lib.ptr(out), stream()), "iga_net_infer")
        if precision == 1:
            packed[(H, W)] = self.weights_version()
        return out

    def backward(self, act, gy, B, H, W):
        """Parameter gradients into ``self.grads``; returns d/d input."""
        import torch

        st = self.struct(H, W)
        L = lib()
        gact = torch.empty_like(act)
        out_off = int(L.iga_net_act_offset(st, B, len(self.arch) - 1))
        gact[out_off: out_off + gy.numel()].copy_(gy.reshape(-1))
        scratch = torch.empty(max(1, int(L.iga_net_scratch_floats(st, B))), dtype=torch.float32,
                              device=act.device)
        _lib.check(L.iga_net_backward(st, B, _lib.ptr(self.params), _lib.ptr(act), _lib.ptr(gact),
                                      _lib.ptr(self.grads), _lib.ptr(scratch), stream()),
                   "iga_net_backward")
        return gact[: B * self.c_in * H * W].view(B, self.c_in, H, W)

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(lib().iga_adam(self.n, _lib.ptr(self.params), _lib.ptr(self.grads), _lib.ptr(self.m),
                                  _lib.ptr(self.v), lr, b1, b2, eps, self.step, stream()), "iga_adam")


def generator(p, seed=0, ks=KS, width=32, device="cuda"):
    """Untrained smoother generator for degree-``p`` operators (its stencils
    are damped Jacobi until it is trained)."""
    return ConvNet(generator_arch((2 * p + 1) ** 2, ks, width), seed=seed, device=device)


# ---------------------------------------------------------------------------
# device helpers
# ---------------------------------------------------------------------------


def features(op: MG.MGOperator, level):
    """``((2p+1)^2, h, w)`` operator-stencil features of one level."""
    import torch

    lv = op.hier.levels[level]
    R = op.hier.p
    if level == 0 and op.grid_ops is not None and R == MG.GRID_RADIUS:
        return op.grid_feats.view(1, -1, lv.h, lv.w)  # written by MGOperator.load
    f = torch.empty(((2 * R + 1) ** 2, lv.n), dtype=torch.float32, device=op.vals64[0].device)
    _lib.check(lib().iga_stencil_features(lv.n, lv.w, R, _lib.ptr(lv.indptr), _lib.ptr(lv.cols),
                                          _lib.ptr(op.vals64[level]), _lib.ptr(f), stream()),
               "iga_stencil_features")
    return f.view(1, -1, lv.h, lv.w)


def csr_f32(n, indptr, cols, nnz, vals, x, mode=0, b=None, out=None, rows_out=None):
    """Batched FP32 CSR product over the rows of ``x (B, n_in)``."""
    import torch

    B = x.shape[0]
    y = torch.empty((B, rows_out or n), dtype=torch.float32, device=x.device) if out is None else out
    _lib.check(lib().iga_csr_f32(rows_out or n, nnz, _lib.ptr(indptr), _lib.ptr(cols), _lib.ptr(vals),
                                 _lib.ptr(x), _lib.ptr(b), _lib.ptr(y), mode, B, x.shape[1], y.shape[1],
                                 stream()), "iga_csr_f32")
    return y


def _op_apply(op, l, x, mode=0, b=None):
    lv = op.hier.levels[l]
    return csr_f32(lv.n, lv.indptr, lv.cols, lv.nnz, op.vals32[l], x, mode, b)


def _transfer(op, l, x, transpose):
    lv = op.hier.levels[l]
    ip, cc, _, v32 = lv.pt_dev if transpose else lv.p_dev
    rows = op.hier.levels[l + 1].n if transpose else lv.n
    return csr_f32(rows, ip, cc, cc.shape[0], v32, x, rows_out=rows)


def stencil_apply(op, l, W, v, transpose=True, accumulate=False, out=None):
    import torch

    lv = op.hier.levels[l]
    y = torch.empty_like(v) if out is None else out
    _lib.check(lib().iga_stencil_apply(lv.h, lv.w, op.ks, _lib.ptr(W), _lib.ptr(op.dsq[l]), _lib.ptr(v),
                                       _lib.ptr(y), int(transpose), int(accumulate), v.shape[0], stream()),
               "iga_stencil_apply")
    return y


# ---------------------------------------------------------------------------
# set-up and solve
# ---------------------------------------------------------------------------


def smoother_stencils(gen: ConvNet, op: MG.MGOperator, levels=LEARNED_LEVELS):
    """Generator output ``(ks*ks, n_l)`` for the ``levels`` finest smoothing
    levels; None (damped Jacobi) below them."""
    out = []
    for l in range(op.hier.n_levels - 1):
        if l < levels:
            # written straight into the operator's stencil buffer of the level
            out.append(gen.infer(features(op, l), op.stencil_buffer(l)))
        else:
            out.append(None)
    return out


def setup(ctx, vals64, gen: ConvNet = None, levels=LEARNED_LEVELS, vals_slot=None):
    """Multigrid operator with the generator's smoothers on the ``levels``
    finest levels and damped Jacobi elsewhere (everywhere when ``gen`` is None).
    ``vals_slot``: level 0's values are read through this one-element int64
    device tensor holding their address (see :meth:`MGOperator.load`)."""
    import torch

    hier = MG.hierarchy(ctx)
    op = MG.operator(hier, omega=OMEGA_COARSE, ks=KS)
    if gen is None or not op.split_load():
        op.load(vals64)
        if gen is not None:
            # the grid cycle's level-0 smoothers read an FP16 copy (the
            # generator's own arithmetic is FP16 on the matrix cores)
            op.set_stencils(smoother_stencils(gen, op, levels), half="copy")
        return op
    # the generator reads only level 0's features: it runs on a second
    # stream beside the Galerkin levels and the dense tail. The launches are
    # enqueued (and captured) in the order the GPU can start them: level 0,
    # the first (longest) Galerkin launch, the generator, the rest -- a graph
    # replay submits its kernels in that order, so the generator's first one
    # is not behind the whole Galerkin chain, nor the chain behind the
    # generator's
    # level 0's CSR copies (read only by the flexible CG's FP64 residual)
    # are written on the generator's stream too, off the Galerkin chain
    op.load(vals64, stage=16, vals_slot=vals_slot)
    main = torch.cuda.current_stream()
    level0 = torch.cuda.Event()
    level0.record(main)
    op.load(vals64, stage=4)
    side = side_stream(op.b.device, "generator")
    side.wait_event(level0)
    with torch.cuda.stream(side):
        stencils = smoother_stencils(gen, op, levels)
        # the grid cycle's FP16 copy, made on the generator's stream
        half = op.write_half(stencils[0])
        op.load(vals64, stage=32, vals_slot=vals_slot)
    op.load(vals64, stage=8)
    main.wait_stream(side)
    op.set_stencils(stencils, half=half)
    return op


class CapturedSetup:
    """The per-operator set-up of :func:`setup` (Galerkin levels, dense tail,
    generator inference on its side stream) captured once in a HIP graph that
    reads the level-0 values through a device slot holding their address: a
    new operator is one 8-byte write and one graph launch, with no copy of
    the values and no host work per kernel."""

    def __init__(self, ctx, gen, levels):
        import torch

        self.gen_ref = weakref.ref(gen)
        dev = ctx.f2f_d.device
        vin = torch.zeros(ctx.assembly.nnz, dtype=torch.float64, device=dev)
        self.slot = torch.full((1,), vin.data_ptr(), dtype=torch.int64, device=dev)
        # one eager run first: lazy buffers stay out of the graph
        self.op = setup(ctx, vin, gen, levels)
        self.stencils = list(self.op.stencils)  # the persistent buffers the graph writes
        torch.cuda.synchronize()
        cap = side_stream(dev, "setup-capture")
        cap.wait_stream(torch.cuda.current_stream())
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=cap):
            setup(ctx, vin, gen, levels, vals_slot=self.slot)
        torch.cuda.synchronize()
        # the eager run packed the FP16 weights, so the captured inference
        # reads them packed (see ConvNet.infer)
        self.packed = gen.weights_version()

    def run(self, vals64):
        # the graph's inference reads the FP16 weights packed when it was
        # captured: pack them again if the parameters have changed since
        gen = self.gen_ref()
        if gen.weights_version() != self.packed:
            lv = self.op.hier.levels[0]
            gen.pack16(lv.h, lv.w)
            self.packed = gen.weights_version()
        # the graph reads vals64 through the slot: it must stay alive (and
        # unchanged) until the set-up has run
        self.slot.fill_(vals64.data_ptr())
        self.graph.replay()
        # the operator may have been set up since without these smoothers
        # (another generator, or none): point it at the graph's buffers again
        self.op.set_stencils(self.stencils, half="ready")
        return self.op


def captured_setup(ctx, vals64, gen: ConvNet, levels=LEARNED_LEVELS):
    """:func:`setup` through the operator's captured set-up graph (captured
    at first use per generator and level count)."""
    op = MG.operator(MG.hierarchy(ctx), omega=OMEGA_COARSE, ks=KS)
    if gen is None or not op.split_load():
        return setup(ctx, vals64, gen, levels)
    graphs = op.__dict__.setdefault("_setup_graphs", {})
    cs = graphs.get(levels)
    if cs is None or cs.gen_ref() is not gen:
        cs = graphs[levels] = CapturedSetup(ctx, gen, levels)
    return cs.run(vals64)


_DEFAULT = {"generator": None}
_SETUP_EVENTS = {}


def use_generator(gen):
    """Register the trained generator whose smoothers
    ``LinearOptions(preconditioner="multigrid")`` uses (None: damped Jacobi)."""
    _DEFAULT["generator"] = gen


def default_generator():
    return _DEFAULT["generator"]


def linear_solve(ctx, vals64, b, opts, adjoint=False):
    """The ``preconditioner="multigrid"`` solve of the Galerkin driver:
    flexible CG with the registered generator's smoothers (or damped Jacobi)
    to ``opts.tol``; ``(u, info)``, counted as a forward or an adjoint solve."""
    LIN.SOLVE_COUNTER["adjoint" if adjoint else "forward"] += 1
    res = solve(ctx, vals64, b.contiguous(), default_generator(), tol=opts.tol,
                maxiter=opts.maxiter if opts.maxiter is not None else 1000)
    return res.x, {"iterations": res.iterations, "fallback_iterations": res.fallback_iterations}


@dataclasses.dataclass
class SolveResult:
    x: object
    iterations: int
    fallback_iterations: int
    relative_residual: float
    seconds: float
    setup_seconds: float


def solve(ctx, vals64, b, gen: ConvNet = None, tol=1e-10, maxiter=1000, graph_iters=8,
          levels=LEARNED_LEVELS, captured=True):
    """Flexible CG with the learned multigrid preconditioner; Jacobi-CG is the
    fallback and the independent FP64 residual the judge. ``captured``: the
    set-up runs from its captured graph (:func:`captured_setup`)."""
    import torch

    from jaxiga_amd import precision as PR

    # the timing events of the set-up, made once per device (creating them
    # is host work in front of the first launch)
    ev = _SETUP_EVENTS.get(b.device)
    if ev is None:
        ev = _SETUP_EVENTS[b.device] = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    op = captured_setup(ctx, vals64, gen, levels) if captured else setup(ctx, vals64, gen, levels)
    ev[1].record()
    # the plan recomputes ||b - K x|| in FP64 from x after the solve
    x, it, rr, bb, rn = op.fcg(b, tol, maxiter, graph_iters, true_residual=True)
    bn = math.sqrt(bb)
    pattern = ctx.assembly
    extra = 0
    if rn > tol * bn:
        r = torch.empty_like(b)
        scratch = torch.empty(2048, dtype=torch.float64, device=b.device)
        rn, bn = PR.csr_residual(pattern, vals64, x, b, r, scratch)
    while rn > tol * bn and extra < 10 * pattern.shape[0]:
        res = LIN.cg_csr(pattern, vals64, r, 0.5 * tol * bn / rn, dinv=op.dinv[0])
        x += res.x
        extra += res.iterations
        rn, bn = PR.csr_residual(pattern, vals64, x, b, r, scratch)
        if res.iterations == 0:
            break
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return SolveResult(x, it, extra, rn / bn if bn > 0 else 0.0, t2 - t0, ev[0].elapsed_time(ev[1]) * 1e-3)


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------


def solver_errors(op: MG.MGOperator, pattern, vals64, b, jacobi_iters=(0, 1, 2, 4, 8, 16, 32, 64),
                  mg_iters=(1, 2, 3)):
    """Errors ``u* - u_k`` (FP64, ``(S, n)``) of truncated Jacobi-CG and
    multigrid-FCG iterates against the converged solution."""
    import torch

    xs = op.fcg(b, 1e-13, 1000)[0]
    out = []
    for k in jacobi_iters:
        xk = torch.zeros_like(b) if k == 0 else LIN.cg_csr(pattern, vals64, b, 1e-30, maxiter=k,
                                                           dinv=op.dinv[0]).x
        out.append(xs - xk)
    for k in mg_iters:
        out.append(xs - op.fcg(b, 1e-30, k, graph_iters=0)[0])
    return torch.stack(out)


def jacobi_errors(op: MG.MGOperator, level, n_start=4, sweeps=(0, 1, 2), seed=0):
    """Errors of damped-Jacobi sweeps ``e <- (I - w D^-1 K) e`` on one level,
    started from random vectors (FP32, ``(n_start * len(sweeps), n_l)``): the
    rough components a smoother must handle."""
    import torch

    lv = op.hier.levels[level]
    g = torch.Generator(device=op.vals32[level].device).manual_seed(seed * 7919 + level)
    E = torch.randn((n_start, lv.n), generator=g, device=g.device, dtype=torch.float32)
    d = op.dsq[level] * op.dsq[level]
    out = []
    for k in range(max(sweeps) + 1):
        if k in sweeps:
            out.append(E.clone())
        E = E - OMEGA0 * d * _op_apply(op, level, E)
    return torch.cat(out)


def level_pairs(op: MG.MGOperator, errors64, n_random=4, seed=0, levels=LEARNED_LEVELS):
    """Per smoothing level (the ``levels`` finest), energy-normalised
    coarse-corrected errors and their residuals ``(E_c, K E_c)`` (FP32,
    ``(S, n_l)``): the solver's errors carried down by restriction, plus
    random-start Jacobi errors of the level."""
    import torch

    pairs = []
    E = errors64.float().contiguous()
    for l in range(min(levels, op.hier.n_levels - 1)):
        El = torch.cat([E, jacobi_errors(op, l, n_random, seed=seed)]) if n_random else E
        R = _op_apply(op, l, El)
        Bc = _transfer(op, l, R, transpose=True)
        Xc = torch.stack([op.vcycle(Bc[s].contiguous(), level=l + 1) for s in range(Bc.shape[0])])
        Ec = El - _transfer(op, l, Xc, transpose=False)
        KEc = _op_apply(op, l, Ec)
        nrm = (Ec.double() * KEc.double()).sum(1).clamp_min(1e-300).sqrt().float()[:, None]
        keep = torch.isfinite(nrm[:, 0]) & (nrm[:, 0] > 0)
        pairs.append(((Ec / nrm)[keep].contiguous(), (KEc / nrm)[keep].contiguous()))
        E = _transfer(op, l, E, transpose=True)
    return pairs


@dataclasses.dataclass
class LevelData:
    """What the training loss needs of one level of one operator."""

    lv: object  # multigrid.LevelStructure
    vals32: object
    dsq: object
    ks: int
    feats: object


class TrainingSet:
    """Per operator and trained level: features, FP32 operator, scalings and
    residual/error pairs (copied, so operators of one space may share a
    multigrid instance)."""

    def __init__(self):
        self.items = []  # [[(LevelData, E_c, R_c) per level]]

    def add(self, ctx, vals64, b, n_random=4, levels=LEARNED_LEVELS):
        op = setup(ctx, vals64, None)
        errs = solver_errors(op, ctx.assembly, vals64, b)
        pairs = level_pairs(op, errs, n_random, seed=len(self.items), levels=levels)
        item = []
        for l, (Ec, Rc) in enumerate(pairs):
            d = LevelData(op.hier.levels[l], op.vals32[l].clone(), op.dsq[l].clone(), op.ks,
                          features(op, l))
            item.append((d, Ec, Rc))
        self.items.append(item)

    @property
    def n_pairs(self):
        return sum(E.shape[0] for item in self.items for _, E, _ in item)


def energy_loss(d: LevelData, W, Ec, Rc):
    """Loss ``mean ||E_c - S^T R_c||_K^2`` of one level and its gradient with
    respect to the stencil weights ``W (ks*ks, n)``."""
    import torch

    lv = d.lv
    L = lib()
    y = torch.empty_like(Rc)
    _lib.check(L.iga_stencil_apply(lv.h, lv.w, d.ks, _lib.ptr(W), _lib.ptr(d.dsq), _lib.ptr(Rc), _lib.ptr(y),
                                   1, 0, Rc.shape[0], stream()), "iga_stencil_apply")
    Z = Ec - y
    KZ = csr_f32(lv.n, lv.indptr, lv.cols, lv.nnz, d.vals32, Z)
    B = Z.shape[0]
    gy = torch.empty_like(Z)
    per = torch.empty(B, dtype=torch.float64, device=Z.device)
    loss = torch.empty(1, dtype=torch.float64, device=Z.device)
    _lib.check(L.iga_energy_loss(B, lv.n, _lib.ptr(Z), _lib.ptr(KZ), _lib.ptr(gy), _lib.ptr(per),
                                 _lib.ptr(loss), stream()), "iga_energy_loss")
    # gy = d loss / d y (y = S^T R_c enters the loss through Z = E_c - y)
    dW = torch.empty_like(W)
    _lib.check(L.iga_stencil_wgrad_t(lv.h, lv.w, d.ks, B, _lib.ptr(d.dsq), _lib.ptr(Rc), _lib.ptr(gy),
                                     _lib.ptr(dW), stream()), "iga_stencil_wgrad_t")
    return loss, dW


def level_data(op: MG.MGOperator, l):
    """:class:`LevelData` view of one level of a loaded multigrid operator."""
    return LevelData(op.hier.levels[l], op.vals32[l], op.dsq[l], op.ks, None)


def train(gen: ConvNet, data: TrainingSet, steps=200, lr=1e-3, batch=None, seed=0, log_every=0):
    """Adam on the mean smoothing loss over (operator, level) pairs of the
    training set -- all operators per step, or ``batch`` of them drawn at
    random; returns the loss history (host floats)."""
    import torch

    rng = np.random.default_rng(seed)
    history = []
    total = torch.zeros_like(gen.grads)
    n_items = len(data.items)
    for step in range(steps):
        total.zero_()
        losses = []
        count = 0
        chosen = range(n_items) if not batch or batch >= n_items else rng.choice(n_items, batch, replace=False)
        for i in chosen:
            for d, Ec, Rc in data.items[i]:
                if Ec.shape[0] == 0:
                    continue
                W4, act = gen.forward(d.feats)
                W = W4[0].reshape(d.ks * d.ks, -1)
                loss, dW = energy_loss(d, W, Ec, Rc)
                gen.backward(act, dW, 1, d.lv.h, d.lv.w)
                total += gen.grads
                losses.append(loss)
                count += 1
        gen.grads.copy_(total / max(count, 1))
        gen.adam(lr)
        history.append(float(torch.cat(losses).mean()))
        if log_every and step % log_every == 0:
            print(f"[train] step {step} loss {history[-1]:.4f}", flush=True)
    return history
