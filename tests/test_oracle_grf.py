"""Pins the random-Fourier-feature field against its defining covariance
sigma^2 exp(-|x-y|^2 / (2 l^2)) (Monte-Carlo over seeds)."""

import numpy as np

from oracle import grf


def test_rff_covariance_matches_squared_exponential():
    ell, sigma = 0.1, 1.5
    pts = np.array([[0.2, 0.3], [0.25, 0.3], [0.2, 0.45], [0.8, 0.9]])
    samples = []
    for seed in range(400):
        om, ph = grf.rff_params(seed, 2, ell, M=256)
        samples.append(grf.rff_field(pts, om, ph, sigma))
    s = np.array(samples)
    cov = s.T @ s / len(s)
    d2 = ((pts[:, None] - pts[None]) ** 2).sum(-1)
    exact = sigma**2 * np.exp(-d2 / (2 * ell**2))
    np.testing.assert_allclose(cov, exact, atol=0.35)
    np.testing.assert_allclose(np.diag(cov), sigma**2, rtol=0.2)


def test_rff_is_deterministic_per_seed():
    a = grf.rff_params(5, 3, 0.2, M=64)
    b = grf.rff_params(5, 3, 0.2, M=64)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
