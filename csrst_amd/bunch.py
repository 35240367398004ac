"""Seeded synthetic electron bunches (counter-based Philox stream).

Positions in metres, momenta as beta*gamma along the given direction,
charges in coulombs (electrons: negative).  The same seed gives the same
arrays on every machine, so the CPU oracle and the GPU receive identical
inputs.
"""
import numpy as np


def gaussian_bunch(n, center, sigma_s, sigma_x, gamma, charge_nC, seed=1, direction=(1.0, 0.0),
                   energy_spread=0.0, chirp=0.0, modulation=None):
    """Gaussian bunch; `chirp` is d(delta)/ds [1/m]; `modulation` = (depth, wavelength) applies a
    sinusoidal density modulation along s by inverse-CDF sampling."""
    rng = np.random.Generator(np.random.Philox(seed))
    d = np.asarray(direction, np.float64)
    d = d / np.hypot(d[0], d[1])
    perp = np.array([-d[1], d[0]])
    s = rng.normal(0.0, sigma_s, n)
    if modulation is not None:
        depth, lam = modulation
        # rejection on the modulated density 1 + depth*sin(2 pi s / lam)
        keep = rng.random(n) * (1.0 + depth) < 1.0 + depth * np.sin(2 * np.pi * s / lam)
        while not keep.all():
            m = (~keep).sum()
            s_new = rng.normal(0.0, sigma_s, m)
            k_new = rng.random(m) * (1.0 + depth) < 1.0 + depth * np.sin(2 * np.pi * s_new / lam)
            s[~keep] = s_new
            idx = np.flatnonzero(~keep)
            keep[idx] = k_new
    x = rng.normal(0.0, sigma_x, n)
    R = np.empty((n, 2))
    R[:, 0] = center[0] + s * d[0] + x * perp[0]
    R[:, 1] = center[1] + s * d[1] + x * perp[1]
    delta = energy_spread * rng.normal(0.0, 1.0, n) + chirp * s
    g = gamma * (1.0 + delta)
    p = np.sqrt(g * g - 1.0)
    P = np.empty((n, 2))
    P[:, 0] = p * d[0]
    P[:, 1] = p * d[1]
    Q = np.full(n, -charge_nC * 1e-9 / n)
    return R, P, Q
