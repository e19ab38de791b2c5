"""Oracle for the learned-preconditioner networks (test infrastructure only).

The reference has no learned preconditioner (``src/jaxiga/solvers/linear.py:
113-116`` allows only ``"none"`` and ``"jacobi"``); the networks are defined by
this project. This module is the CPU FP64 statement of that definition --
forward pass and exact gradients of networks of 3x3 convolutions with optional
transposed-convolution + skip layers, per-channel biases and ReLUs -- against
which the GPU forward/backward kernels are checked. It is itself pinned by
torch-CPU ``conv2d``/``conv_transpose2d`` autograd and by finite differences
(``tests/test_oracle_learned.py``).

Conventions (shared with ``jaxiga_amd.learned``): tensors are NCHW; a
stride-``s`` 3x3 convolution with zero padding 1 maps ``H -> (H - 1)//s + 1``
and has weights ``(C_out, C_in, 3, 3)``; the transposed convolution has
weights ``(C_in, C_out, 3, 3)`` and an explicit output size, and is the exact
adjoint of the forward convolution with the same weights. A layer computes
``act(conv(x) [+ skip] + bias)``. Layers are
``(name, kind, c_in, c_out, stride, input, skip, bias, relu)``; parameters
are a dict of weights ``name`` and biases ``name + ".b"``.
"""

from __future__ import annotations

import numpy as np


def generator_arch(c_in, ks=7, width=32):
    """The smoother generator: operator stencil features -> ks*ks stencil weights."""
    return (
        ("g0", "conv", c_in, width, 1, "x", None, True, True),
        ("g1", "conv", width, width, 1, "g0", None, True, True),
        ("w", "conv", width, ks * ks, 1, "g1", None, True, False),
    )


# a 4-level U-Net exercising strided and transposed layers, skips, biases and ReLUs
UNET_ARCH = (
    ("a0", "conv", 2, 8, 1, "x", None, True, True),
    ("a1", "conv", 8, 16, 2, "a0", None, False, True),
    ("a2", "conv", 16, 16, 2, "a1", None, True, False),
    ("a3", "conv", 16, 16, 2, "a2", None, True, True),
    ("b3", "conv", 16, 16, 1, "a3", None, False, False),
    ("u2", "up", 16, 16, 2, "b3", "a2", True, True),
    ("u1", "up", 16, 16, 2, "u2", "a1", False, True),
    ("u0", "up", 16, 8, 2, "u1", "a0", True, False),
    ("y", "conv", 8, 3, 1, "u0", None, True, False),
)


def weight_shape(layer):
    _, kind, cin, cout = layer[:4]
    return (cout, cin, 3, 3) if kind == "conv" else (cin, cout, 3, 3)


def n_params(arch):
    return int(sum(np.prod(weight_shape(l)) + (l[3] if l[7] else 0) for l in arch))


def out_size(h, s):
    return (h - 1) // s + 1


def conv(x, w, s):
    """Forward 3x3 convolution, padding 1, stride s."""
    B, C, H, W = x.shape
    Ho, Wo = out_size(H, s), out_size(W, s)
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    out = np.zeros((B, w.shape[0], Ho, Wo))
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky : ky + s * (Ho - 1) + 1 : s, kx : kx + s * (Wo - 1) + 1 : s]
            out += np.einsum("oc,bchw->bohw", w[:, :, ky, kx], patch)
    return out


def conv_t(x, w, s, H, W):
    """Transposed convolution (adjoint of ``conv`` with weights ``w`` viewed as
    (C_out_fwd=C_in_t, C_in_fwd=C_out_t)), output size (H, W)."""
    B, C, h, wd = x.shape
    out = np.zeros((B, w.shape[1], H + 2, W + 2))
    for ky in range(3):
        for kx in range(3):
            contrib = np.einsum("co,bchw->bohw", w[:, :, ky, kx], x)
            out[:, :, ky : ky + s * (h - 1) + 1 : s, kx : kx + s * (wd - 1) + 1 : s] += contrib
    return out[:, :, 1:-1, 1:-1]


def conv_wgrad(x, gy, s):
    """d loss / d w for ``y = conv(x, w, s)`` given ``gy``; shape (C_out, C_in, 3, 3)."""
    B, C, H, W = x.shape
    Ho, Wo = gy.shape[2:]
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    g = np.zeros((gy.shape[1], C, 3, 3))
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky : ky + s * (Ho - 1) + 1 : s, kx : kx + s * (Wo - 1) + 1 : s]
            g[:, :, ky, kx] = np.einsum("bohw,bchw->oc", gy, patch)
    return g


def forward(arch, params, x):
    """Returns the output and the (post-activation) output of every layer."""
    act = {"x": x}
    for name, kind, cin, cout, s, src, skip, bias, relu in arch:
        inp = act[src]
        if kind == "conv":
            out = conv(inp, params[name], s)
        else:
            H, W = act[skip].shape[2:]
            out = conv_t(inp, params[name], s, H, W) + act[skip]
        if bias:
            out = out + params[name + ".b"][None, :, None, None]
        if relu:
            out = np.maximum(out, 0.0)
        act[name] = out
    return act[arch[-1][0]], act


def backward(arch, params, act, gy):
    """Gradients w.r.t. every parameter and the input, given d loss / d output."""
    grads = {}
    gact = {arch[-1][0]: gy}
    for name, kind, cin, cout, s, src, skip, bias, relu in reversed(arch):
        g = gact.pop(name)
        if relu:
            g = g * (act[name] > 0)
        if bias:
            grads[name + ".b"] = g.sum((0, 2, 3))
        inp = act[src]
        if kind == "conv":
            grads[name] = conv_wgrad(inp, g, s)
            gin = conv_t(g, params[name], s, *inp.shape[2:])
        else:
            # out = conv_t(inp, w) + skip ; conv_t is the adjoint of conv, so
            # d/dw[c, o, k] = sum_y inp[c, y] g[o, s*y + k - 1]
            grads[name] = conv_wgrad(g, inp, s)
            gin = conv(g, params[name], s)
            gact[skip] = gact.get(skip, 0) + g
        gact[src] = gact.get(src, 0) + gin
    return grads, gact["x"]


def init_params(arch, seed=0):
    rng = np.random.default_rng(seed)
    params = {}
    for layer in arch:
        shape = weight_shape(layer)
        params[layer[0]] = rng.normal(size=shape) * np.sqrt(2.0 / (layer[2] * 9))
        if layer[7]:
            params[layer[0] + ".b"] = rng.normal(size=layer[3]) * 0.1
    return params
