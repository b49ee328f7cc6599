"""Host logic of the sweep driver and the autotuners that needs no GPU:
sweep point sampling, the droplet end-state classifier, the Laplace
pressure probe, and the candidate launches of the batch and single-domain
autotuners."""
import numpy as np
import pytest


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


def test_batch_candidates_grids():
    """Every shape with 4, 8 and 16 blocks per CU, plus one block per work
    item where the batch has more items than the largest of those grids;
    the two-steps-per-launch tiles (kinds 4 and 5) only with the E4 force."""
    from idmx import tune
    shapes = tune.batch_shapes()
    assert len(set(shapes)) == len(shapes) == 32
    for k in (4, 5):
        assert sorted(s for s in shapes if s[0] == k) == [(k, 16, 16), (k, 32, 8), (k, 32, 16)]
    assert not [s for s in tune.batch_shapes('E8') if s[0] in (4, 5)]
    base = tune.batch_candidates(256)
    assert len(base) == 3 * len(shapes)
    assert {c[3] for c in base} == {1024, 2048, 4096}
    items = {s: (5000 if s[0] in (1, 3, 4, 5) else 4096) for s in shapes}
    cands = tune.batch_candidates(256, items)
    extra = [c for c in cands if c not in base]
    assert sorted(extra) == sorted(s + (5000,) for s in shapes if s[0] in (1, 3, 4, 5))


def test_droplet_centre_offsets():
    from idmx import configs, sweep
    for c, want in (((32.0, 32.0), (0.0, 0.0)), ((31.5, 32.5), (-0.5, 0.5)),
                    ((33.0, 30.0), (1.0, -2.0))):
        n = configs.droplet((64, 64), 12.0, 2.0, centre=c)
        got = sweep.droplet_centre(n)
        assert abs(got[0] - want[0]) < 1e-6 and abs(got[1] - want[1]) < 1e-6, (c, got)


def test_run_sweep_rejects_unknown_rest_init():
    from idmx import sweep
    with pytest.raises(ValueError):
        sweep.run_sweep(None, [sweep.SweepPoint(L=16, G=-1.6, R=4.0)], rest_init='hot')


def test_sim_candidates_cover_every_variant():
    """The single-domain tuner times every launch variant, the x-shifted
    storage (10) included, on every tile and march length."""
    from idmx import tune
    for D, tiles in ((2, ((64, 1), (128, 1), (256, 1))),
                     (3, ((32, 8), (64, 4), (128, 2), (256, 2)))):
        cands = tune.sim_candidates(D, 256)
        assert len(cands) == len(set(cands)) == len(tiles) * 3 * 11
        assert {c[3] for c in cands} == set(range(11))
        assert {c[:2] for c in cands if c[3] == 10} == set(tiles)
    assert {c[2] for c in tune.sim_candidates(3, 40)} == {32}
    assert {c[2] for c in tune.sim_candidates(2, 17)} == {17}
