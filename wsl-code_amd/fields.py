"""Synthetic inputs: seeded Gaussian random fields for log-permeability."""
import numpy as np


def grf_theta(x, y, sigma=1.0, ell=0.05, seed=0, features=512, chunk=65536):
    """Sample of a zero-mean Gaussian random field with squared-exponential
    covariance sigma^2 exp(-r^2 / (2 ell^2)) at the points (x, y), by random
    Fourier features: theta = sigma sqrt(2/M) sum_m cos(w_m . x + b_m) with
    w_m ~ N(0, I / ell^2) and b_m ~ U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((features, 2)) / ell
    b = rng.uniform(0.0, 2.0 * np.pi, features)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    out = np.empty(len(x))
    for s in range(0, len(x), chunk):
        ph = np.outer(x[s:s + chunk], w[:, 0]) + np.outer(y[s:s + chunk], w[:, 1]) + b
        out[s:s + chunk] = np.cos(ph).sum(1)
    return sigma * np.sqrt(2.0 / features) * out
