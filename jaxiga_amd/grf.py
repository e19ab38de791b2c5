"""Seeded Gaussian random fields (synthetic coefficient inputs).

The reference has no random-field sampler (its field-valued coefficients
arrive through ``PointField``, ``src/jaxiga/methods/_common.py:28-51``). The
synthetic coefficients ``a = exp(g)`` of this build use random Fourier
features for a squared-exponential covariance
``sigma^2 exp(-|x - y|^2 / (2 l^2))``:

    g(x) = sigma sqrt(2/M) sum_m cos(omega_m . x + b_m),
    omega_m ~ N(0, l^-2 I)  (drawn first, shape (M, dim)),  b_m ~ U[0, 2 pi),

with ``numpy.random.default_rng(seed)``; the draws are made on the host and
the field is evaluated in FP64 on the GPU (``iga_rff_eval``).
"""

from __future__ import annotations

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd.runtime import lib, stream


def rff_params(seed, dim, ell, M=4096):
    rng = np.random.default_rng(seed)
    omega = rng.normal(size=(M, dim)) / float(ell)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=M)
    return omega, phase


def rff_eval(points, omega, phase, sigma):
    """``g`` at device points ``(..., dim)`` (float64 tensor)."""
    import torch

    dev = points.device
    pts = points.reshape(-1, points.shape[-1]).contiguous()
    om = torch.as_tensor(omega, dtype=torch.float64, device=dev).contiguous()
    ph = torch.as_tensor(phase, dtype=torch.float64, device=dev).contiguous()
    out = torch.empty(pts.shape[0], dtype=torch.float64, device=dev)
    amp = float(sigma) * np.sqrt(2.0 / len(phase))
    _lib.check(lib().iga_rff_eval(pts.shape[0], pts.shape[1], _lib.ptr(pts), len(phase), _lib.ptr(om),
                                  _lib.ptr(ph), amp, _lib.ptr(out), stream()), "iga_rff_eval")
    return out.reshape(points.shape[:-1])


def gauss_point_field(ctx, seed, ell, sigma, M=4096):
    """``g`` at the interior Gauss points of a method context, ``(n_e, n_q)``."""
    omega, phase = rff_params(seed, ctx.space.dim, ell, M)
    return rff_eval(ctx.gauss_points(), omega, phase, sigma)


def poisson_instance(refine, seed, ell=0.1, sigma=1.5, degree=3, source_ell=0.25, source_sigma=1.0,
                     space=None):
    """A seeded variable-coefficient Poisson problem on the unit square:
    ``-div(a grad u) = f``, ``u = 0`` on the boundary, ``a = exp(g)`` with
    ``g`` a GRF (``ell``, ``sigma``) sampled at the Gauss points and ``f`` a
    smooth random source (a second GRF). The patch is the bilinear unit square
    elevated to ``degree`` and refined to ``2**refine`` elements per side
    (or ``space``, reused with its cached patterns). Returns ``(ctx, vals,
    rhs)``: the method context, FP64 CSR values and the free load vector,
    assembled on the GPU."""
    import jaxiga_amd as jx
    from jaxiga_amd import methods as M

    V = space if space is not None else jx.FunctionSpace(
        jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(degree).refine(refine))
    om_s, ph_s = rff_params(10_000 + seed, 2, source_ell)

    def source(x, p=None):
        return rff_eval(x.T.contiguous(), om_s, ph_s, source_sigma)

    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: np.ones(len(x), bool))],
                      source=source)
    ctx = M.build_context(prob)
    a = gauss_point_field(ctx, seed, ell, sigma)
    a.exp_()
    vals, rhs, _ = M.assemble(ctx, {"a0": jx.PointField(a)})
    return ctx, vals, rhs
