"""Seeded Gaussian random fields by random Fourier features (input synthesis).

The reference has no random-field sampler; the synthetic coefficient
``a(x) = exp(g(x))`` is defined by this project: ``g`` is a zero-mean field
with squared-exponential covariance ``sigma^2 exp(-|x-y|^2 / (2 l^2))``,
approximated with ``M`` random Fourier features

    g(x) = sigma * sqrt(2/M) * sum_m cos(omega_m . x + b_m),
    omega_m ~ N(0, l^-2 I),  b_m ~ U[0, 2 pi),

drawn from ``numpy.random.default_rng(seed)`` in that order (omega first,
shape ``(M, dim)``, then b). This module is the CPU statement of that
definition; the GPU evaluates the same draws.
"""

from __future__ import annotations

import numpy as np


def rff_params(seed, dim, ell, M=4096):
    rng = np.random.default_rng(seed)
    omega = rng.normal(size=(M, dim)) / float(ell)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=M)
    return omega, phase


def rff_field(points, omega, phase, sigma):
    """``g`` at ``points (n, dim)``, FP64, chunked over points."""
    points = np.asarray(points, float)
    M = len(phase)
    out = np.empty(len(points))
    step = max(1, 2_000_000 // M)
    for lo in range(0, len(points), step):
        arg = points[lo : lo + step] @ omega.T + phase
        out[lo : lo + step] = np.cos(arg).sum(1)
    return sigma * np.sqrt(2.0 / M) * out
