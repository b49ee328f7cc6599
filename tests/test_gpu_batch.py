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
    for shape in [(0, 128, 32, 0), (0, 64, 16, 512), (0, 256, 64, 0), (1, 32, 8, 0),
                  (1, 64, 4, 300), (1, 16, 16, 0), (1, 32, 4, 0), (2, 128, 32, 0),
                  (2, 64, 8, 300), (3, 32, 8, 0), (3, 16, 16, 0)]:
        print('launch shape', shape, flush=True)
        b.set_launch(*shape)
        for i, n0 in enumerate(starts):
            b.set_density(i, n0)
        b.run(150)   # one graph of 100 steps, then 50 single steps
        for i, p in enumerate(pts):
            assert rel_err(b.get_density(i).ravel(), refs[i]) <= tol, (shape, i, p)
    b.close()


def test_batch_one_item_per_workgroup(tenet):
    """Grids of one work item per workgroup and beyond (the autotuner's
    largest candidates; surplus workgroups exit at once), and the same grid
    after domains leave the live set, when it exceeds the item count."""
    from idmx.sweep import Batch, cold_start
    pts = _points()
    b = Batch(tenet, pts, 'fp64')
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    refs = [_single(tenet, p, starts[i], 150, 'fp64') for i, p in enumerate(pts)]
    for kind, tx, H in [(3, 32, 8), (1, 64, 4), (2, 128, 8)]:
        b.set_launch(kind, tx, H)
        items = b.get_launch()[4]
        assert items > 0
        for grid in (items, 4 * items):
            print('launch shape', (kind, tx, H, grid), flush=True)
            b.set_launch(kind, tx, H, grid)
            for i, n0 in enumerate(starts):
                b.set_density(i, n0)
            b.run(150)
            for i, p in enumerate(pts):
                assert rel_err(b.get_density(i).ravel(), refs[i]) <= 1e-13, (kind, grid, i)
    b.set_launch(3, 32, 8)
    full = b.get_launch()[4]
    b.set_launch(3, 32, 8, full)
    for i, n0 in enumerate(starts):
        b.set_density(i, n0)
    b.run(100)
    b.set_active([0, 1, 0, 1])
    kind, tx, H, grid, items = b.get_launch()
    assert grid == full and 0 < items < full
    b.run(50)
    for i in (1, 3):
        assert rel_err(b.get_density(i).ravel(), refs[i]) <= 1e-13, i
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


@pytest.mark.parametrize('es', ['E4', 'E6', 'E8'])
def test_batch_tiles_match_oracle_odd_sizes(tenet, es):
    """Tiles wider and taller than the domain, ragged last tiles and the
    R = 2 halo ring of E6 / E8, against the oracle's Meta step."""
    from oracle import cpuref
    from idmx.sweep import Batch, SweepPoint, cold_start
    pts = [SweepPoint(L=70, G=-1.6, R=15.0), SweepPoint(L=9, G=-5.0 / 3, R=2.5),
           SweepPoint(L=37, G=-1.5, R=9.0)]
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    refs = []
    for p, n0 in zip(pts, starts):
        st = cpuref.State('D2Q9', (p.L, p.L))
        cpuref.init_pop(st, n0, np.zeros(2 * p.L * p.L))
        cpuref.meta_run(st, 'D2' + es, p.G, p.tau, nsteps=40)
        refs.append(st.pop.reshape(9, -1).sum(axis=0))
    b = Batch(tenet, pts, 'fp64', estencil=es)
    for shape in [(1, 32, 8, 0), (1, 64, 4, 0), (1, 16, 16, 7), (1, 32, 4, 0), (0, 64, 8, 0),
                  (3, 32, 8, 0), (3, 64, 4, 5), (2, 64, 8, 0), (2, 128, 4, 0)]:
        print('es', es, 'launch shape', shape, flush=True)
        b.set_launch(*shape)
        for i, n0 in enumerate(starts):
            b.set_density(i, n0)
        b.run(40)
        for i, p in enumerate(pts):
            assert rel_err(b.get_density(i).ravel(), refs[i]) <= 1e-12, (shape, i, p)
    b.close()


def test_sweep_fp32_then_fp64_converges(tenet):
    """FP32 phase to its rounding floor, then FP64 (from the symmetric rest
    state of the FP32 density) to tol; cold starts are never reset, so a
    non-monotone residual cannot loop."""
    from idmx.sweep import SweepPoint, run_sweep
    pts = [SweepPoint(L=64, G=-5.0 / 3, R=14.0), SweepPoint(L=48, G=-1.55, R=12.0)]
    # march items while both domains live, tiles once the 64^2 one has
    # converged (the end-of-sweep launch switch)
    res, fields = run_sweep(tenet, pts, mode='fp32+fp64', tol=1e-9, check_every=200,
                            max_steps=40000, launch=(0, 128, 16, 0),
                            launch_small=((1, 32, 8, 0), 64 * 64),
                            launch32=(0, 64, 8, 0), launch32_small=((1, 16, 16, 0), 64 * 64))
    print('steps', res.steps, 'history tail', res.history[-3:], flush=True)
    assert res.converged.all() and not res.fallback.any()
    ref, fref = run_sweep(tenet, pts, mode='fp64', tol=1e-9, check_every=200, max_steps=40000)
    assert ref.converged.all()
    from idmx.sweep import laplace_dp
    for f, g, p in zip(fields, fref, pts):
        # the FP32 phase may translate the droplet by rounding-level momentum,
        # which moves the field but not the pressure jump
        print('max |n32+64 - n64|', np.abs(f - g).max(), flush=True)
        # the hand-over restores the start mass, so both runs relax to the
        # same symmetric steady state
        dp, dp_ref = laplace_dp(f, p.G), laplace_dp(g, p.G)
        assert abs(dp - dp_ref) <= 1e-6 * abs(dp_ref), (p, dp, dp_ref)
        assert np.abs(f - g).max() < 1e-6


def test_symmetrize_projects_and_keeps_mass(tenet):
    """The projection of an asymmetric field is invariant under the domain's
    mirror images (and the transpose when square), keeps the total mass,
    leaves a symmetric field unchanged, and a target mass shifts it
    uniformly."""
    from idmx.sweep import Batch, SweepPoint, cold_start
    pts = [SweepPoint(L=40, G=-1.6, R=9.0), SweepPoint(L=33, G=-1.6, R=8.0)]
    b = Batch(tenet, pts, 'fp64')
    rng = np.random.default_rng(5)
    for i, p in enumerate(pts):
        n0 = cold_start(p).reshape(p.L, p.L) + 0.01 * rng.random((p.L, p.L))
        b.set_density(i, n0)
        b.symmetrize(i)
        ns = b.get_density(i)
        c = p.L // 2
        idx = (2 * c - np.arange(p.L)) % p.L
        assert np.allclose(ns, ns[:, idx], atol=1e-14, rtol=0)
        assert np.allclose(ns, ns[idx, :], atol=1e-14, rtol=0)
        assert np.allclose(ns, ns.T, atol=1e-14, rtol=0)
        assert abs(ns.sum() - n0.sum()) <= 1e-12 * n0.sum()
        b.symmetrize(i)
        assert np.abs(b.get_density(i) - ns).max() <= 1e-14
        target = np.zeros(len(pts))
        target[i] = n0.sum() + 3.0
        b.symmetrize(i, mass=target)
        nm = b.get_density(i)
        assert abs(nm.sum() - target[i]) <= 1e-12 * target[i]
        assert np.abs(nm - ns - 3.0 / nm.size).max() <= 1e-13
    b.close()
