"""The batched sweep kernel against the single-domain fused step, the
per-domain convergence check, graph replay and the sweep driver's fallback."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _points():
    from idmx.sweep import SweepPoint
    return [SweepPoint(L=64, G=-5.0 / 3, R=12.0), SweepPoint(L=96, G=-1.5, R=30.0),
            SweepPoint(L=160, G=-1.75, R=20.0), SweepPoint(L=70, G=-1.6, R=15.0)]


def _single(tenet, p, n0, steps, precision):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    eng = SCEngine(tenet, 'D2Q9', 'D2E4', (p.L, p.L), 'exp', p.G, p.tau, precision=precision,
                   rho0=1.0)
    n = IdpyMemory.OnDevice(n0, tenet)
    u = IdpyMemory.Zeros(2 * p.L * p.L, np.float64, tenet)
    a, b = PopArray(eng, tenet), PopArray(eng, tenet)
    eng.init_pop(a, n, u)
    if eng.run(a, b, steps):
        a, b = b, a
    f = eng.pop_to_host_soa(a)
    eng.close()
    return f.sum(axis=0)


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-13), ('fp32s', 1e-6)])
def test_batch_matches_single_domain(tenet, precision, tol):
    from idmx.sweep import Batch, cold_start
    pts = _points()
    b = Batch(tenet, pts, precision)
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    refs = [_single(tenet, p, starts[i], 150, precision) for i, p in enumerate(pts)]
    for shape in [(128, 32, 0), (64, 16, 512), (256, 64, 0)]:
        print('launch shape', shape, flush=True)
        b.set_launch(*shape)
        for i, n0 in enumerate(starts):
            b.set_density(i, n0)
        b.run(150)   # one graph of 100 steps, then 50 single steps
        for i, p in enumerate(pts):
            assert rel_err(b.get_density(i).ravel(), refs[i]) <= tol, (shape, i, p)
    b.close()


def test_batch_check_and_active_set(tenet):
    from idmx.sweep import Batch, cold_start
    pts = _points()
    b = Batch(tenet, pts, 'fp64')
    for i, p in enumerate(pts):
        b.set_density(i, cold_start(p))
    b.run(20)
    d1 = b.check()
    assert np.all(d1 > 0)
    frozen = b.get_density(1).copy()
    b.set_active([1, 0, 1, 1])
    b.run(40)
    d2 = b.check()
    assert d2[1] == 0.0 and np.all(d2[[0, 2, 3]] > 0)
    assert np.array_equal(b.get_density(1), frozen)
    b.close()


def test_sweep_converges_and_falls_back(tenet):
    from idmx.sweep import SweepPoint, run_sweep
    pts = [SweepPoint(L=64, G=-5.0 / 3, R=14.0), SweepPoint(L=64, G=-1.55, R=18.0)]
    bad = [np.full(64 * 64, np.nan), None]
    from idmx.sweep import cold_start
    bad[1] = cold_start(pts[1])
    res, fields = run_sweep(tenet, pts, start=bad, mode='fp64', tol=1e-7, check_every=200,
                            max_steps=60000)
    assert res.converged.all()
    assert res.fallback[0] and not res.fallback[1]
    assert all(np.isfinite(f).all() for f in fields)
