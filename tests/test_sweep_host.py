"""Host logic of the sweep driver that needs no GPU: sweep point sampling,
the droplet end-state classifier and the Laplace pressure probe."""
import numpy as np


def test_sample_points_ranges():
    from idmx.sweep import sample_points
    pts = sample_points(200, 7, Ls=tuple(range(64, 257, 32)))
    assert all(p.L in range(64, 257, 32) for p in pts)
    assert all(-5.5 / 3 <= p.G <= -4.2 / 3 for p in pts)
    assert all(0.1 * p.L <= p.R <= 0.35 * p.L for p in pts)
    assert sample_points(5, 7)[0] == sample_points(5, 7)[0]


def test_droplet_classifier():
    from idmx.sweep import SweepPoint, cold_start, is_centred_droplet
    for L, R in [(64, 10.0), (96, 25.0), (160, 40.0)]:
        assert is_centred_droplet(cold_start(SweepPoint(L=L, G=-5 / 3, R=R)))
    L = 96
    x = np.abs(np.arange(L) - L // 2)
    stripe = np.where(x[None, :] < 20, 1.9, 0.12) * np.ones((L, 1))
    assert not is_centred_droplet(stripe)
    bubble = 2.02 - cold_start(SweepPoint(L=L, G=-5 / 3, R=20.0))
    assert not is_centred_droplet(bubble)
    shifted = np.roll(cold_start(SweepPoint(L=L, G=-5 / 3, R=20.0)).reshape(L, L), 12, axis=1)
    assert not is_centred_droplet(shifted)


def test_laplace_probe_sign_and_value():
    from idmx.sweep import SweepPoint, cold_start, laplace_dp
    G = -5 / 3
    n = cold_start(SweepPoint(L=64, G=G, R=12.0)).reshape(64, 64)
    p = lambda v: v / 3 + 0.5 * G * (1 - np.exp(-v)) ** 2
    assert abs(laplace_dp(n, G) - (p(n[32, 32]) - p(n[-1, -1]))) < 1e-15
