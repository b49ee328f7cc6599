"""Device observables and the sweep's stopping rule against numpy on the same
device fields: max |a - b| (the steady-state check), the velocity check of
CheckUConvergenceSCMP (sum |u - old_u|, max u^2; idpy/LBM/
MultiPhaseLoopChecks.py:96-163), the batch's per-domain max |dn|, and the
batch state of domains left out of the live set across odd step counts."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def test_max_abs_diff_matches_numpy(tenet):
    from idmx.IdpyCode import IdpyMemory
    rng = np.random.default_rng(11)
    for n in (1, 255, 256, 1000, 65537, 3 * 10 ** 6 + 17):
        a = rng.standard_normal(n)
        b = a + rng.standard_normal(n) * 10.0 ** rng.uniform(-14, 0, n)
        da, db = IdpyMemory.OnDevice(a, tenet), IdpyMemory.OnDevice(b, tenet)
        # a maximum is exact: no rounding in max |a - b|
        assert IdpyMemory.MaxAbsDiff(da, db) == np.abs(a - b).max(), n
        b[n // 2] = np.nan
        db.H2D(b)
        assert IdpyMemory.MaxAbsDiff(da, db) == np.inf, n
        da.Free()
        db.Free()


@pytest.mark.parametrize('D', [2, 3])
def test_check_u_matches_numpy(tenet, D):
    from idmx import _lib
    from idmx.IdpyCode import IdpyMemory
    rng = np.random.default_rng(12 + D)
    V = 70001
    u = rng.uniform(-0.05, 0.05, D * V)
    old = u + rng.uniform(-1e-6, 1e-6, D * V)
    du, dold = IdpyMemory.OnDevice(u, tenet), IdpyMemory.OnDevice(old, tenet)
    s, m = C.c_double(), C.c_double()
    _lib.call('idmx_check_u', tenet.handle, du.ptr, dold.ptr, V, D, C.byref(s), C.byref(m))
    us, os_ = u.reshape(D, V), old.reshape(D, V)
    want_s = np.abs(us - os_).sum()
    want_m = (us * us).sum(axis=0).max()
    assert abs(s.value - want_s) <= 1e-13 * want_s
    # the per-site sum u_x^2 + u_y^2 (+ u_z^2) is formed in the same order
    assert abs(m.value - want_m) <= 1e-15 * want_m
    assert np.array_equal(dold.D2H(), u)   # old_u <- u
    du.Free()
    dold.Free()


def _pts():
    from idmx.sweep import SweepPoint
    return [SweepPoint(L=64, G=-5.0 / 3, R=12.0), SweepPoint(L=96, G=-1.5, R=30.0),
            SweepPoint(L=70, G=-1.6, R=15.0)]


def test_batch_check_matches_numpy(tenet):
    """idmx_batch_check returns, per live domain, max |n - n_prev| of the
    densities at this and the previous check, exactly as numpy computes it
    from the fields read back, and 0 for a domain out of the live set."""
    from idmx.sweep import Batch, cold_start
    pts = _pts()
    b = Batch(tenet, pts, 'fp64')
    for i, p in enumerate(pts):
        b.set_density(i, cold_start(p) + 1e-3 * np.sin(np.arange(p.L * p.L)))
    b.run(30)
    b.check()
    before = [b.get_density(i) for i in range(len(pts))]
    for steps in (10, 200, 7):
        b.run(steps)
        now = [b.get_density(i) for i in range(len(pts))]
        got = b.check()
        for i in range(len(pts)):
            assert got[i] == np.abs(now[i] - before[i]).max(), (steps, i)
        before = now
    b.set_active([1, 0, 1])
    b.run(20)
    got = b.check()
    assert got[1] == 0.0
    now = [b.get_density(i) for i in range(len(pts))]
    for i in (0, 2):
        assert got[i] == np.abs(now[i] - before[i]).max(), i
    b.close()


@pytest.mark.parametrize('launch', [(1, 32, 8, 0), (3, 32, 8, 0), (2, 64, 8, 0)])
def test_batch_odd_steps_keep_inactive_domain(tenet, launch):
    """A domain out of the live set keeps its state across runs of odd step
    counts (single steps after the graph replays flip the buffers), and the
    live domains follow the single-domain step."""
    from idmx.sweep import Batch, cold_start
    from test_gpu_batch import _single
    pts = _pts()
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    total = 41 + 1 + 101 + 3
    refs = [_single(tenet, p, starts[i], total, 'fp64') for i, p in enumerate(pts)]
    b = Batch(tenet, pts, 'fp64')
    b.set_launch(*launch)
    for i, n0 in enumerate(starts):
        b.set_density(i, n0)
    b.set_active([1, 0, 1])
    frozen = b.get_density(1).copy()
    for steps in (41, 1, 101, 3):
        print('launch', launch, 'run', steps, flush=True)
        b.run(steps)
        assert np.array_equal(b.get_density(1), frozen), steps
    for i in (0, 2):
        assert rel_err(b.get_density(i).ravel(), refs[i]) <= 1e-13, i
    # back in the live set, it continues from its own start
    b.set_active([1, 1, 1])
    b.run(total)
    assert rel_err(b.get_density(1).ravel(), refs[1]) <= 1e-13
    b.close()
