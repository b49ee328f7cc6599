"""Shared helpers of the parity tests: run the oracle and the HIP path on the
same seeded inputs and compare field by field."""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def smooth_field(dims, seed, lo=0.3, hi=1.8):
    """A smooth random density (sum of a few periodic modes) in [lo, hi]."""
    rng = np.random.default_rng(seed)
    axes = np.meshgrid(*[np.arange(L) / L for L in dims], indexing='ij')
    axes = [a.transpose(tuple(range(len(dims)))[::-1]) for a in axes]
    f = np.zeros(axes[0].shape)
    for _ in range(4):
        k = rng.integers(1, 3, size=len(dims))
        ph = rng.uniform(0, 2 * np.pi, size=len(dims))
        term = np.ones_like(f)
        for a, kk, p in zip(axes, k, ph):
            term = term * np.cos(2 * np.pi * kk * a + p)
        f += rng.uniform(0.5, 1.0) * term
    f = (f - f.min()) / (f.max() - f.min())
    return (lo + (hi - lo) * f).ravel()


def droplet_field(dims, seed, noise=1e-3):
    """Tanh droplet (n_g = 0.1, n_l = 2.0, width 2, R = min(dims)/4) at a
    seeded off-centre position plus uniform noise: near coexistence, so
    it stays stable at the workloads' coupling."""
    from idmx import configs
    rng = np.random.default_rng(seed)
    c = [L / 2 + rng.uniform(-2, 2) for L in dims]
    n = configs.droplet(dims, min(dims) / 4, 2.0, centre=c)
    return n + rng.uniform(-noise, noise, size=n.shape)
