"""Oracle for the learned preconditioner network (test infrastructure only).

The reference has no learned preconditioner (``src/jaxiga/solvers/linear.py:
113-116`` allows only ``"none"`` and ``"jacobi"``); the network is defined by
this project. This module is the CPU FP64 statement of that definition --
forward pass and exact gradients of a 4-level, bias-free, activation-free
convolutional U-Net on the control-point grid -- against which the GPU
forward/backward kernels are checked. It is itself pinned by torch-CPU
``conv2d``/``conv_transpose2d`` autograd and by finite differences
(``tests/test_oracle_learned.py``).

Conventions (shared with ``jaxiga_amd.learned``): tensors are NCHW; a
stride-``s`` 3x3 convolution with zero padding 1 maps ``H -> (H - 1)//s + 1``
and has weights ``(C_out, C_in, 3, 3)``; the transposed convolution has
weights ``(C_in, C_out, 3, 3)`` and an explicit output size, and is the exact
adjoint of the forward convolution with the same weights.
"""

from __future__ import annotations

import numpy as np

# (name, kind, c_in, c_out, stride, input, skip) ; kind: "conv" | "up"
ARCH = (
    ("a0", "conv", 1, 8, 1, "x", None),
    ("a1", "conv", 8, 16, 2, "a0", None),
    ("a2", "conv", 16, 32, 2, "a1", None),
    ("a3", "conv", 32, 64, 2, "a2", None),
    ("b3", "conv", 64, 64, 1, "a3", None),
    ("c3", "conv", 64, 64, 1, "b3", None),
    ("d3", "conv", 64, 64, 1, "c3", None),
    ("u2", "up", 64, 32, 2, "d3", "a2"),
    ("b2", "conv", 32, 32, 1, "u2", None),
    ("c2", "conv", 32, 32, 1, "b2", None),
    ("u1", "up", 32, 16, 2, "c2", "a1"),
    ("b1", "conv", 16, 16, 1, "u1", None),
    ("u0", "up", 16, 8, 2, "b1", "a0"),
    ("y", "conv", 8, 1, 1, "u0", None),
)


def weight_shape(layer):
    _, kind, cin, cout, _, _, _ = layer
    return (cout, cin, 3, 3) if kind == "conv" else (cin, cout, 3, 3)


def n_params():
    return int(sum(np.prod(weight_shape(l)) for l in ARCH))


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


def forward(params, x):
    """Returns the output and the activations of every layer."""
    act = {"x": x}
    for layer in ARCH:
        name, kind, cin, cout, s, src, skip = layer
        inp = act[src]
        if kind == "conv":
            out = conv(inp, params[name], s)
        else:
            H, W = act[skip].shape[2:]
            out = conv_t(inp, params[name], s, H, W) + act[skip]
        act[name] = out
    return act["y"], act


def backward(params, act, gy):
    """Gradients w.r.t. every weight and the input, given d loss / d y."""
    grads = {}
    gact = {"y": gy}
    for layer in reversed(ARCH):
        name, kind, cin, cout, s, src, skip = layer
        g = gact.pop(name)
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


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    params = {}
    for layer in ARCH:
        shape = weight_shape(layer)
        fan = layer[2] * 9
        params[layer[0]] = rng.normal(size=shape) * np.sqrt(1.0 / fan)
    return params


def normalized_mse(pred, target):
    """Per-sample ||pred - target||^2 / ||target||^2, averaged; and its gradient."""
    B = pred.shape[0]
    d = pred - target
    den = (target**2).reshape(B, -1).sum(1)
    loss = float(((d**2).reshape(B, -1).sum(1) / den).mean())
    grad = 2.0 * d / den[:, None, None, None] / B
    return loss, grad
