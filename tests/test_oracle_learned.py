"""Pins the network oracle: forward against torch-CPU conv2d/conv_transpose2d
in FP64, gradients against torch autograd and central finite differences."""

import numpy as np
import torch
import torch.nn.functional as F

from oracle import learned as L


def _torch_forward(params, x):
    act = {"x": x}
    for name, kind, cin, cout, s, src, skip in L.ARCH:
        w = params[name]
        if kind == "conv":
            act[name] = F.conv2d(act[src], w, stride=s, padding=1)
        else:
            H, W = act[skip].shape[2:]
            h, wd = act[src].shape[2:]
            op = (H - ((h - 1) * s - 2 + 3), W - ((wd - 1) * s - 2 + 3))
            act[name] = F.conv_transpose2d(act[src], w, stride=s, padding=1, output_padding=op) + act[skip]
    return act["y"]


def test_forward_and_gradients_match_torch_autograd():
    rng = np.random.default_rng(1)
    params = L.init_params(3)
    x = rng.normal(size=(2, 1, 19, 21))
    t = rng.normal(size=(2, 1, 19, 21))
    y, act = L.forward(params, x)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in params.items()}
    tx = torch.tensor(x, requires_grad=True)
    ty = _torch_forward(tp, tx)
    np.testing.assert_allclose(y, ty.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss, gy = L.normalized_mse(y, t)
    grads, gx = L.backward(params, act, gy)
    tt = torch.tensor(t)
    tl = (((ty - tt) ** 2).sum((1, 2, 3)) / (tt**2).sum((1, 2, 3))).mean()
    tl.backward()
    assert abs(loss - tl.item()) < 1e-12
    for k in params:
        np.testing.assert_allclose(grads[k], tp[k].grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(gx, tx.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_gradient_finite_differences():
    rng = np.random.default_rng(2)
    params = L.init_params(4)
    x = rng.normal(size=(1, 1, 11, 9))
    t = rng.normal(size=(1, 1, 11, 9))
    y, act = L.forward(params, x)
    _, gy = L.normalized_mse(y, t)
    grads, _ = L.backward(params, act, gy)
    for name in ("a0", "d3", "u1", "y"):
        idx = tuple(rng.integers(0, n) for n in params[name].shape)
        h = 1e-6
        p1 = {k: v.copy() for k, v in params.items()}
        p2 = {k: v.copy() for k, v in params.items()}
        p1[name][idx] += h
        p2[name][idx] -= h
        fd = (L.normalized_mse(L.forward(p1, x)[0], t)[0] - L.normalized_mse(L.forward(p2, x)[0], t)[0]) / (2 * h)
        assert abs(fd - grads[name][idx]) <= 1e-6 * max(1.0, abs(fd))


def test_parameter_count():
    assert 150_000 < L.n_params() < 250_000
