"""The batched tiles advanced two steps per launch (kind 4: the first step's
populations kept in LDS for the tile plus a 2-site ring; kind 5: the same on
the x-shifted storage between the launches of a replayed graph) against the
one-step tiles (kind 1) and the oracle: the same arithmetic in the same
order, so the densities (and, kind 5 against kind 4, the populations) agree
bit for bit; odd step counts, domains smaller than a tile, ragged last
tiles, persistent grids of fewer blocks than items (down to one), inactive
domains, FP64 and FP32-shifted."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _points():
    from idmx.sweep import SweepPoint
    return [SweepPoint(L=70, G=-1.6, R=15.0), SweepPoint(L=9, G=-5.0 / 3, R=2.5),
            SweepPoint(L=37, G=-1.5, R=9.0), SweepPoint(L=160, G=-1.75, R=30.0)]


@pytest.mark.parametrize('precision', ['fp64', 'fp32s'])
def test_two_steps_per_launch_equals_one_step_tiles(tenet, precision):
    from idmx.sweep import Batch, cold_start
    pts = _points()
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    b = Batch(tenet, pts, precision)
    ref = {}
    for nsteps in (150, 41, 1, 2):
        b.set_launch(1, 32, 8, 0)
        for i, n0 in enumerate(starts):
            b.set_density(i, n0)
        b.run(nsteps)
        ref[nsteps] = [b.get_density(i) for i in range(len(pts))]
    items = {}
    for shape in [(4, 32, 8, 0), (4, 16, 16, 0), (4, 32, 16, 0), (4, 32, 8, 7), (4, 16, 16, 1000),
                  (4, 32, 16, 13), (4, 16, 16, 1), (5, 32, 8, 0), (5, 16, 16, 0), (5, 32, 16, 13),
                  (5, 16, 16, 1)]:
        for nsteps in (150, 41, 1, 2):
            b.set_launch(*shape)
            for i, n0 in enumerate(starts):
                b.set_density(i, n0)
            b.run(nsteps)
            for i in range(len(pts)):
                got = b.get_density(i)
                assert np.array_equal(got, ref[nsteps][i]), \
                    (shape, nsteps, i, rel_err(got.ravel(), ref[nsteps][i].ravel()))
        print(precision, 'shape', shape, 'identical to kind 1', flush=True)
    b.close()


@pytest.mark.parametrize('launch', [(4, 32, 16, 0), (4, 32, 8, 1024), (5, 32, 8, 0),
                                    (5, 16, 16, 1024)])
def test_two_steps_keep_inactive_domains_and_match_oracle(tenet, launch):
    from oracle import cpuref
    from idmx.sweep import Batch, cold_start
    pts = _points()
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    b = Batch(tenet, pts, 'fp64')
    b.set_launch(*launch)
    for i, n0 in enumerate(starts):
        b.set_density(i, n0)
    b.set_active([1, 1, 0, 1])
    before = b.get_density(2)
    b.run(41)
    assert np.array_equal(b.get_density(2), before)
    b.set_active([1, 1, 1, 1])
    for i, n0 in enumerate(starts):
        b.set_density(i, n0)
    b.run(40)
    for i, (p, n0) in enumerate(zip(pts, starts)):
        st = cpuref.State('D2Q9', (p.L, p.L))
        cpuref.init_pop(st, n0, np.zeros(2 * p.L * p.L))
        cpuref.meta_run(st, 'D2E4', p.G, p.tau, nsteps=40)
        err = rel_err(b.get_density(i).ravel(), st.pop.reshape(9, -1).sum(axis=0))
        print('kind', launch[0], 'vs oracle, domain', i, 'L', p.L, err, flush=True)
        assert err <= 1e-12
    b.close()


def test_kind4_refused_without_e4(tenet):
    from idmx import _lib
    from idmx.sweep import Batch
    b = Batch(tenet, _points()[:2], 'fp64', estencil='E8')
    for kind in (4, 5):
        with pytest.raises(_lib.IdmxError):
            b.set_launch(kind, 32, 8, 0)
    b.close()


@pytest.mark.parametrize('precision', ['fp64', 'fp32s'])
def test_xshift_kind5_populations_equal_kind4(tenet, precision):
    """Kind 5 leaves the streamed storage after every run: its populations
    equal kind 4's bit for bit after runs that replay the graph (100 steps
    per replay), leave steps over, or are too short for a graph."""
    from idmx.sweep import Batch, cold_start
    pts = _points()
    starts = [cold_start(p) + 1e-4 * np.cos(np.arange(p.L * p.L)) for p in pts]
    b = Batch(tenet, pts, precision)
    out = {}
    for kind in (4, 5):
        b.set_launch(kind, 32, 8, 0)
        for i, n0 in enumerate(starts):
            b.set_density(i, n0)
        res = []
        for nsteps in (204, 3, 6, 101):
            b.run(nsteps)
            res.append([b.get_populations(i) for i in range(len(pts))])
        out[kind] = res
    for r4, r5 in zip(out[4], out[5]):
        for i in range(len(pts)):
            assert np.array_equal(r4[i], r5[i]), (precision, i)
    b.close()
