"""Pins the network oracle: forward against torch-CPU conv2d/conv_transpose2d
in FP64 (with biases and ReLUs), gradients against torch autograd and central
finite differences, for the smoother generator and a 4-level U-Net."""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import learned as L

ARCHS = {"generator": L.generator_arch(49, 7, 16), "unet": L.UNET_ARCH}


def _torch_forward(arch, params, x):
    act = {"x": x}
    for name, kind, cin, cout, s, src, skip, bias, relu in arch:
        w = params[name]
        if kind == "conv":
            out = F.conv2d(act[src], w, stride=s, padding=1)
        else:
            H, W = act[skip].shape[2:]
            h, wd = act[src].shape[2:]
            op = (H - ((h - 1) * s - 2 + 3), W - ((wd - 1) * s - 2 + 3))
            out = F.conv_transpose2d(act[src], w, stride=s, padding=1, output_padding=op) + act[skip]
        if bias:
            out = out + params[name + ".b"][None, :, None, None]
        if relu:
            out = torch.relu(out)
        act[name] = out
    return act[arch[-1][0]]


def _loss(y, t):
    return float(((y - t) ** 2).sum())


@pytest.mark.parametrize("kind", sorted(ARCHS))
def test_forward_and_gradients_match_torch_autograd(kind):
    arch = ARCHS[kind]
    rng = np.random.default_rng(1)
    params = L.init_params(arch, 3)
    x = rng.normal(size=(2, arch[0][2], 19, 21))
    y, act = L.forward(arch, params, x)
    t = rng.normal(size=y.shape)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in params.items()}
    tx = torch.tensor(x, requires_grad=True)
    ty = _torch_forward(arch, tp, tx)
    np.testing.assert_allclose(y, ty.detach().numpy(), rtol=1e-12, atol=1e-12)
    grads, gx = L.backward(arch, params, act, 2.0 * (y - t))
    ((ty - torch.tensor(t)) ** 2).sum().backward()
    for k in params:
        np.testing.assert_allclose(grads[k], tp[k].grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(gx, tx.grad.numpy(), rtol=1e-10, atol=1e-10)


def test_gradient_finite_differences():
    arch = L.UNET_ARCH
    rng = np.random.default_rng(2)
    params = L.init_params(arch, 4)
    x = rng.normal(size=(1, 2, 11, 9))
    y, act = L.forward(arch, params, x)
    t = rng.normal(size=y.shape)
    grads, _ = L.backward(arch, params, act, 2.0 * (y - t))
    for name in ("a0", "a0.b", "b3", "u1", "u0.b", "y"):
        idx = tuple(rng.integers(0, n) for n in params[name].shape)
        h = 1e-6
        p1 = {k: v.copy() for k, v in params.items()}
        p2 = {k: v.copy() for k, v in params.items()}
        p1[name][idx] += h
        p2[name][idx] -= h
        fd = (_loss(L.forward(arch, p1, x)[0], t) - _loss(L.forward(arch, p2, x)[0], t)) / (2 * h)
        assert abs(fd - grads[name][idx]) <= 1e-5 * max(1.0, abs(fd)), name


def test_parameter_counts():
    assert L.n_params(L.generator_arch(49, 7, 32)) == 49 * 32 * 9 + 32 + 32 * 32 * 9 + 32 + 32 * 49 * 9 + 49
