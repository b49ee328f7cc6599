"""ProfileModel (the lattice-resolution part of the learned warm start) on
synthetic droplets drawn from a known law: tanh profiles whose liquid and
vapour densities and width depend on (G, R) through the model's own
features, plus a fixed radial bump. The fit must recover the law (the
rendered start equals the true field), describe() must return a pure tanh's
parameters, and predict() must hit the requested mass."""
import numpy as np
import pytest

from idmx import warmstart as ws
from idmx.sweep import SweepPoint


def _law(G, R):
    g, rho = (G + 1.6) / 0.25, 10.0 / R
    n_l = 2.0 + 0.3 * g + 0.02 * rho
    n_g = 0.1 - 0.03 * g + 0.004 * rho
    W = 1.8 - 0.2 * g
    return n_l, n_g, W


def _bump(s):
    return 0.01 * s * np.exp(-s * s)


def _field(G, L, R):
    n_l, n_g, W = _law(G, R)
    s = (ws._centre_dist(L) - R) / W
    return n_g + (n_l - n_g) * (0.5 * (1 - np.tanh(s)) + _bump(s))


def _points(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.choice((64, 96, 128)))
        out.append(SweepPoint(L=L, G=float(rng.uniform(-1.8, -1.45)),
                              R=float(rng.uniform(0.2 * L, 0.33 * L))))
    return out


def test_describe_pure_tanh():
    L, R = 128, 30.3
    n_l, n_g, W = _law(-1.7, R)
    s = (ws._centre_dist(L) - R) / W
    n = n_g + (n_l - n_g) * 0.5 * (1 - np.tanh(s))
    d = ws.ProfileModel.describe(n, L)
    assert abs(d[2] - R) < 1e-9 and abs(d[3] - W) < 1e-9
    assert abs(d[0] - n_l) < 1e-3 * (n_l - n_g) and abs(d[1] - n_g) < 1e-6
    assert np.nanmax(np.abs(d[4])) < 2e-3


def test_describe_rejects_non_droplets():
    L = 64
    assert ws.ProfileModel.describe(np.full(L * L, 0.7), L) is None
    n = _field(-1.7, L, 20.0)
    assert ws.ProfileModel.describe(n.max() + n.min() - n, L) is None   # a bubble


def test_fit_recovers_the_law():
    pts = _points(60, 1)
    pm = ws.ProfileModel()
    assert pm.fit(pts, [_field(p.G, p.L, p.R) for p in pts]) == 60
    for p in _points(8, 2):
        true = _field(p.G, p.L, p.R)
        start = pm.render(p.G, p.L, p.R)
        assert np.abs(start - true).max() < 2e-3 * true.max(), p


def test_predict_hits_mass_and_radius():
    pts = _points(60, 3)
    pm = ws.ProfileModel()
    pm.fit(pts, [_field(p.G, p.L, p.R) for p in pts])
    ev = _points(6, 4)
    masses = [_field(p.G, p.L, p.R).sum() for p in ev]
    starts = pm.predict(ev, masses)
    for p, m, st in zip(ev, masses, starts):
        assert st is not None
        assert abs(st.sum() - m) < 1e-9 * m
        # the mass fixes the radius: the start is the true field
        assert np.abs(st - _field(p.G, p.L, p.R)).max() < 5e-3


def test_predict_declines_impossible_mass():
    pts = _points(40, 5)
    pm = ws.ProfileModel()
    pm.fit(pts, [_field(p.G, p.L, p.R) for p in pts])
    p = SweepPoint(L=64, G=-1.6, R=15.0)
    assert pm.predict([p], [0.05 * 64 * 64])[0] is None      # below the vapour density
    assert pm.predict([p], [2.5 * 64 * 64])[0] is None       # above the liquid density


def test_unfitted_model_refuses():
    pm = ws.ProfileModel()
    with pytest.raises(RuntimeError):
        pm.predict([SweepPoint(L=64, G=-1.6, R=15.0)], [1000.0])
    with pytest.raises(RuntimeError):
        pm.render(-1.6, 64, 15.0)
    assert pm.p_uniform(-1.6, 0.3, 64) == 0.0


def test_fit_needs_enough_fields():
    pts = _points(5, 6)
    with pytest.raises(ValueError):
        ws.ProfileModel().fit(pts, [_field(p.G, p.L, p.R) for p in pts])


def test_uniform_classifier():
    """Droplets at strong coupling / high mean density, uniform gas at weak
    coupling / low mean density: the classifier separates them, a uniform
    point starts exactly uniform at its mass, and a droplet point still gets
    the profile."""
    rng = np.random.default_rng(8)
    pts, fields = [], []
    for _ in range(80):
        L = int(rng.choice((64, 96, 128)))
        if rng.random() < 0.5:
            p = SweepPoint(L=L, G=float(rng.uniform(-1.8, -1.6)), R=float(rng.uniform(0.2 * L, 0.33 * L)))
            fields.append(_field(p.G, p.L, p.R))
        else:
            p = SweepPoint(L=L, G=float(rng.uniform(-1.5, -1.4)), R=float(rng.uniform(0.1, 0.15) * L))
            fields.append(np.full(L * L, float(rng.uniform(0.15, 0.3))))
        pts.append(p)
    pm = ws.ProfileModel()
    pm.fit(pts, fields)
    assert pm.cls_coef is not None and pm.n_uniform >= 20
    gas = SweepPoint(L=96, G=-1.45, R=12.0)
    m_gas = 0.2 * 96 * 96
    st = pm.predict([gas], [m_gas])[0]
    assert pm.is_uniform(gas, m_gas)
    assert np.ptp(st) == 0.0 and abs(st.sum() - m_gas) < 1e-9 * m_gas
    drop = SweepPoint(L=96, G=-1.7, R=25.0)
    m_drop = _field(drop.G, drop.L, drop.R).sum()
    assert not pm.is_uniform(drop, m_drop)
    assert np.ptp(pm.predict([drop], [m_drop])[0]) > 1.0
