"""Autotuner for the launch shapes of the fused step and the batched sweep.

The fused step's shape is (tile width, planes per block march); the batched
sweep's is (x-chunk width TX, rows per item H, persistent grid size). The
tuner times every candidate on scratch buffers (the caller's state is not
touched), keeps the fastest and caches it per device and workload signature
in a JSON file ($IDMX_TUNE_CACHE, default ~/.cache/idmx_tune.json).
"""
import json
import os

import numpy as np

_CACHE = os.environ.get('IDMX_TUNE_CACHE', os.path.join(os.path.expanduser('~'), '.cache',
                                                         'idmx_tune.json'))


def _load():
    try:
        return json.load(open(_CACHE))
    except Exception:
        return {}


def _store(key, val):
    c = _load()
    c[key] = val
    try:
        os.makedirs(os.path.dirname(_CACHE), exist_ok=True)
        json.dump(c, open(_CACHE, 'w'), indent=1)
    except OSError:
        pass


def sim_candidates(D, L_march):
    tiles = ((64, 1), (128, 1), (256, 1)) if D == 2 else ((32, 8), (64, 4), (128, 2), (256, 2))
    Hs = [h for h in (32, 64, 128) if h <= L_march] or [L_march]
    return [(tx, ty, h, var) for tx, ty in tiles for h in Hs
            for var in (1, 0, 2, 3, 4, 5, 6, 7, 8, 9, 10)]


def tune_sim(engine, n0=None, steps=10, use_cache=True, apply=True):
    """Times the fused step of an SCEngine for each (tile, H, hist)
    candidate on scratch populations initialised from n0 (default: uniform
    1.0 plus small noise) and sets the fastest on the engine. Returns (best,
    timings in ms per step)."""
    from .IdpyCode import IdpyMemory
    from .LBM._base import PopArray
    key = 'sim|%s|%s|%s|%s|%s|%s' % (engine.tenet.GetDeviceName(), engine.xi_name,
                                      engine.e_name, 'x'.join(map(str, engine.dims)),
                                      engine.precision, 'mrt' if engine.is_mrt else 'bgk')
    if use_cache:
        c = _load().get(key)
        if c:
            if apply:
                engine.set_launch(*c['best'])
            return tuple(c['best']), c['timings']
    t = engine.tenet
    V, D = engine.V, engine.D
    if n0 is None:
        n0 = 1.0 + 1e-3 * np.sin(np.arange(V) * 0.37)
    n = IdpyMemory.OnDevice(np.asarray(n0, dtype=np.float64), t)
    u = IdpyMemory.Zeros(V * D, np.float64, t)
    a, b = PopArray(engine, t), PopArray(engine, t)
    engine.init_pop(a, n, u)
    timings = {}
    old = engine.get_launch()
    for cand in sim_candidates(D, engine.dims[-1]):
        if not engine.launch_ok(*cand):
            continue
        engine.set_launch(*cand)
        engine.run(a, b, 2)
        t_short = engine.last_run_ms()
        engine.run(a, b, 2 + steps)
        # difference of two run lengths: the per-run costs (the pull form's
        # two conversion passes) cancel, so every variant is timed per step
        timings[','.join(map(str, cand))] = max(engine.last_run_ms() - t_short, 1e-6) / steps
    best = min(timings, key=timings.get)
    best = tuple(int(v) for v in best.split(','))
    engine.set_launch(*(best if apply else old))
    for x in (a, b, n, u):
        x.Free()
    _store(key, {'best': best, 'timings': timings})
    return best, timings


def batch_shapes(estencil='E4'):
    """(kind, tx, H): marching items and tiles, push and pull form, and with
    the E4 force tiles advanced two steps per launch (kind 4), also on the
    x-shifted storage (kind 5)."""
    shapes = [(k, tx, H) for k in (0, 2) for tx in (64, 128, 256) for H in (8, 16, 32)]
    shapes += [(k, tx, H) for k in (1, 3) for tx, H in ((64, 4), (32, 8), (16, 16), (32, 4))]
    if estencil == 'E4':
        shapes += [(k, tx, H) for k in (4, 5) for tx, H in ((32, 8), (16, 16), (32, 16))]
    return shapes


def batch_candidates(cus, items=None, estencil='E4'):
    """(kind, tx, H, grid): every shape with three persistent grid sizes,
    plus, where items maps a shape to its work-item count on the batch being
    tuned, a grid of one item per workgroup. One item per workgroup measured
    4 % faster than the best persistent grid for the tiles of a full
    128-domain D2Q9 FP64 batch (3.8 M sites); the grid stays as tuned when
    domains finish, and the surplus workgroups exit at once."""
    out = [s + (g,) for s in batch_shapes(estencil) for g in (cus * 4, cus * 8, cus * 16)]
    if items:
        out += [s + (items[s],) for s in batch_shapes(estencil)
                if items.get(s, 0) > cus * 16]
    return out


def tune_batch(tenet, points, precision, estencil='E4', steps=20, use_cache=True):
    """Times the batched step for every (kind, tx, H, grid) candidate on a
    scratch batch of the same domains (cold-start droplets) and returns the
    fastest plus all timings (ms per step)."""
    from .sweep import Batch, cold_start
    sig = '%d:%d:%d' % (len(points), sum(p.L * p.L for p in points), max(p.L for p in points))
    key = 'batch4|%s|%s|%s|%s' % (tenet.GetDeviceName(), precision, estencil, sig)
    if use_cache:
        c = _load().get(key)
        if c:
            return tuple(c['best']), c['timings']
    b = Batch(tenet, points, precision, estencil)
    for i, p in enumerate(points):
        b.set_density(i, cold_start(p))
    cus = b.get_launch()[3] // 8
    items = {}
    for s in batch_shapes(estencil):
        b.set_launch(*s, cus * 4)
        items[s] = b.get_launch()[4]
    timings = {}
    for cand in batch_candidates(cus, items, estencil):
        b.set_launch(*cand)
        # each length runs twice and the second run is timed: the first
        # captures the step graph for that length, whose capture and
        # instantiation (tens of microseconds, more than a small batch's
        # step) are not step time; the difference of the two lengths
        # cancels the pull form's conversions
        b.run(steps)
        t_a = b.run(steps)
        b.run(2 * steps)
        t_b = b.run(2 * steps)
        timings[','.join(map(str, cand))] = max(t_b - t_a, 1e-6) / steps
    b.close()
    best = min(timings, key=timings.get)
    best = tuple(int(v) for v in best.split(','))
    _store(key, {'best': best, 'timings': timings})
    return best, timings


def tune_batch_small(tenet, points, precision, estencil='E4', count=None, steps=20):
    """The launch for the second part of a sweep, once the live domains hold
    no more sites than the `count` largest domains of `points` (default: the
    larger half, at least 3), tuned on those domains. Returns ((kind, tx, H,
    grid), their total sites), the form run_sweep's launch_small takes.
    Measured on a 128-point sweep: switching at the larger half instead of
    the 3 largest cut the batched wall time from 10.1 s to 8.8 s (the tile
    launches beat the marching ones once a few dozen domains are left)."""
    if count is None:
        count = max(3, len(points) // 2)
    big = sorted(points, key=lambda p: p.L, reverse=True)[:count]
    best, _ = tune_batch(tenet, big, precision, estencil, steps=steps, use_cache=False)
    return best, sum(p.L * p.L for p in big)
