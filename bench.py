#!/usr/bin/env python3
"""Benchmark of the fused Shan-Chen step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload (the metric's D3Q19 case): D3Q19 MRT Shan-Chen, 256^3
periodic, FP64 populations, E4 force, psi = 1 - exp(-n), one spherical droplet
R = 40 (n_l = 2.0, n_g = 0.1, tanh width 2), synthetic initial field.
Other workloads: d2q9_4096_e8 (64 droplets, D2Q9 4096^2 FP64 E8 BGK),
d3q19_bgk_256 (same as default with BGK), d3q19_mrt_256_fp32s (FP32 shifted
populations).

Prints ONE JSON line. value = MLUPS = V * K / (wall time of the K timed steps)
with the populations resident in HBM. roofline.achieved = algorithmic bytes
per launch (2 Q sizeof(pop) per site: every population read once and written
once) / the kernel's average duration from HIP events recorded on the
library's stream around the K launches. cpu_baseline = the oracle's serial
C++ restatement of the reference CTypes step on one host core, timed on a
bounded sample (see "sample").
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'idea.deploy_amd'))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0

# fused-step variants (idmx_sc_set_launch)
VARIANTS = {0: 'push, re-read', 1: 'push, registers', 2: 'push, re-read + prefetch',
            3: 'push, re-read, >=3 waves/SIMD', 4: 'push, re-read + prefetch, >=3 waves/SIMD',
            5: 'pull, registers', 6: 'pull, re-read', 7: 'push, re-read, 2 planes/barrier',
            8: 'pull, re-read, 2 planes/barrier',
            9: 'push, registers, lead plane copied to LDS by DMA',
            10: 'variant 9 on the x-shifted storage (line-aligned push stores)'}

WORKLOADS = {
    'd3q19_mrt_256': dict(xi='D3Q19', es='D3E4', dims=(256, 256, 256), coll='mrt',
                          prec='fp64', init='sphere3d'),
    'd3q19_bgk_256': dict(xi='D3Q19', es='D3E4', dims=(256, 256, 256), coll='bgk',
                          prec='fp64', init='sphere3d'),
    'd3q19_mrt_256_fp32s': dict(xi='D3Q19', es='D3E4', dims=(256, 256, 256), coll='mrt',
                                prec='fp32s', init='sphere3d'),
    'd2q9_4096_e8': dict(xi='D2Q9', es='D2E8', dims=(4096, 4096), coll='bgk', prec='fp64',
                         init='droplets2d'),
}


def mrt_S(xi):
    """Relaxation rates of the MRT workloads: omega = 1/tau (tau = 1) on the
    second-order (shear and bulk) moments and 1 on all others."""
    from idmx.LBM import hermite
    _, tuples = hermite.invertible_set(xi)
    return np.diag([1.0 for _ in tuples])


def workload_desc(name, w):
    d = {'workload': name, 'lattice': w['xi'], 'dims': list(w['dims']),
         'force_stencil': w['es'], 'collision': w['coll'], 'psi': '1 - exp(-n)',
         'SC_G': -5.0 / 3.0, 'tau': 1.0, 'populations': w['prec']}
    if w['init'] == 'sphere3d':
        d['init'] = 'sphere R=40, n_l=2.0, n_g=0.1, tanh width 2'
    else:
        d['init'] = '64 seeded droplets R in [10,60], n_l=2.0, n_g=0.1'
    return d


def run_gpu(name, steps, warmup, tune=True, launch=None, two_pass=False):
    from idmx import configs
    from idmx.IdpyCode import IdpyMemory, Tenet
    from idmx.LBM import hermite
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    w = WORKLOADS[name]
    tenet = Tenet(0)
    n0 = configs.sphere3d() if w['init'] == 'sphere3d' else configs.droplets2d()
    mrt = hermite.mrt_kernel_tables(w['xi'], mrt_S(w['xi'])) if w['coll'] == 'mrt' else None
    rho0 = 1.0 if w['prec'] == 'fp32s' else 0.0
    eng = SCEngine(tenet, w['xi'], w['es'], w['dims'], 'exp', configs.SC_G, 1.0,
                   precision=w['prec'], rho0=rho0, mrt=mrt)
    V, D = eng.V, eng.D
    if launch:
        eng.set_launch(*launch)
    elif tune and not two_pass:
        from idmx import tune as tuner
        tuner.tune_sim(eng, steps=5, use_cache=False)
    n = IdpyMemory.OnDevice(n0, tenet)
    u = IdpyMemory.Zeros(V * D, np.float64, tenet)
    psi = IdpyMemory.Zeros(V, np.float64, tenet) if two_pass else None
    pop = PopArray(eng, tenet)
    swap = PopArray(eng, tenet)
    eng.init_pop(pop, n, u)
    del n0
    m0, _ = eng.mass_momentum(pop)
    kw = {'n': n, 'psi': psi, 'two_pass': True} if two_pass else {}
    if warmup > 0 and eng.run(pop, swap, warmup, **kw):
        pop, swap = swap, pop
    tenet.Sync()
    t0 = time.perf_counter()
    if eng.run(pop, swap, steps, **kw):
        pop, swap = swap, pop
    tenet.Sync()
    wall = time.perf_counter() - t0
    ev_ms = eng.last_run_ms()
    m1, _ = eng.mass_momentum(pop)
    launch = eng.get_launch()
    itemsize = 8 if w['prec'] == 'fp64' else 4
    B_alg = 2 * eng.Q * itemsize
    res = {
        'V': V, 'Q': eng.Q, 'wall_s': wall, 'event_ms': ev_ms,
        'mlups': V * steps / wall / 1e6,
        'kernel_ms': ev_ms / steps,
        'B_alg': B_alg,
        'mass_rel_drift': abs(m1 - m0) / m0,
        'launch': {'tile_x': launch[0], 'tile_y': launch[1], 'planes_per_block': launch[2],
                   'variant': launch[3], 'variant_name': VARIANTS.get(launch[3], '?')},
        'device': tenet.GetDeviceName(),
    }
    eng.close()
    tenet.End()
    return res


def cpu_baseline(name, budget_s=12.0):
    """Oracle (serial C++ restatement of the reference's CTypes step, AoS
    layout as on CTypes) on one core, on a bounded sample of the workload:
    the same lattice kind, stencil, collision and initial profile on a
    smaller periodic box, timed over whole steps."""
    from idmx import configs
    from oracle import cpuref, hermite_mrt
    w = WORKLOADS[name]
    dims = (64, 64, 64) if w['xi'] == 'D3Q19' else (512, 512)
    if w['init'] == 'sphere3d':
        n0 = configs.droplet(dims, 10.0, 2.0)
    else:
        n0 = configs.droplets2d(L=512, count=4, seed=configs.SEED0 + 2)
    st = cpuref.State(w['xi'], dims, aos=True)
    V = int(np.prod(dims))
    tmp = cpuref.State(w['xi'], dims)
    cpuref.init_pop(tmp, n0, np.zeros(V * len(dims)))
    st.pop[:] = tmp.pop.reshape(st.Q, V).T.ravel()
    mats = None
    if w['coll'] == 'mrt':
        import sympy as sp
        mats = hermite_mrt.mrt_matrices(w['xi'], sp.eye(st.Q))
    t0 = time.perf_counter()
    cpuref.meta_run(st, w['es'], configs.SC_G, 1.0, nsteps=1, mrt=mats)
    one = time.perf_counter() - t0
    k = max(1, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    cpuref.meta_run(st, w['es'], configs.SC_G, 1.0, nsteps=k, mrt=mats)
    dt = time.perf_counter() - t0
    out = {'value': V * k / dt / 1e6, 'unit': 'MLUPS', 'cores': 1, 'kind': 'port',
           'sample': '%s %s %s FP64 (AoS, as CTypes) on %s, %d steps in %.1f s'
                     % (w['xi'], w['es'], w['coll'].upper(), 'x'.join(map(str, dims)), k, dt),
           'cpu_model': cpu_model()}
    # the other physics of the workloads, same serial oracle, a few seconds each
    out['other_samples'] = {
        'd2q9_e4_128_droplet': _cpu_sample('D2Q9', 'D2E4', (128, 128),
                                           configs.droplet2d_noise(), 4.0),
        'd2q9_e8_512_droplets': _cpu_sample('D2Q9', 'D2E8', (512, 512),
                                            configs.droplets2d(L=512, count=4,
                                                               seed=configs.SEED0 + 2), 4.0),
    }
    return out


def device_clocks():
    """The GPU's current clock levels as rocm-smi reports them (read only;
    the same kernel has measured up to 15 % apart on different boxes), or
    None where rocm-smi is not available."""
    import subprocess
    try:
        p = subprocess.run(['rocm-smi', '--showclocks', '--json'], capture_output=True,
                           text=True, timeout=20)
        d = json.loads(p.stdout)
    except Exception:
        return None
    out = {}
    for card, v in d.items():
        if isinstance(v, dict):
            out[card] = {k: val for k, val in v.items() if 'clock' in k.lower() or 'clk' in k.lower()}
    return out or None


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its core count."""
    name = '?'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                name = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return '%s (%d logical CPUs visible)' % (name, os.cpu_count() or 0)


def _cpu_sample(xi, es, dims, n0, budget_s):
    """MLUPS of the oracle's BGK step (AoS, one core) on a bounded sample."""
    from idmx import configs
    from oracle import cpuref
    V = int(np.prod(dims))
    st = cpuref.State(xi, dims, aos=True)
    tmp = cpuref.State(xi, dims)
    cpuref.init_pop(tmp, n0, np.zeros(V * len(dims)))
    st.pop[:] = tmp.pop.reshape(st.Q, V).T.ravel()
    t0 = time.perf_counter()
    cpuref.meta_run(st, es, configs.SC_G, 1.0, nsteps=1)
    k = max(1, int(budget_s / max(time.perf_counter() - t0, 1e-6)))
    t0 = time.perf_counter()
    cpuref.meta_run(st, es, configs.SC_G, 1.0, nsteps=k)
    dt = time.perf_counter() - t0
    return {'mlups': round(V * k / dt / 1e6, 3), 'steps': k, 'seconds': round(dt, 2),
            'lattice': 'x'.join(map(str, dims)), 'stencil': es}


def run_sweep_study(n_train, n_eval, Lmax, epochs, log=print, train_tol=1e-8,
                    train_max_steps=400000, ablations=True):
    """Time to steady state (|dn|_max < 1e-10 between checks 200 steps apart)
    of a (G, R, L) sweep of centred droplets: plain per-point FP64 runs from
    the cold start vs batched runs (FP64, and FP32 then FP64) from the cold
    start and from the learned warm start. L uniform in {64, 96, ..., Lmax},
    G uniform in [-5.5, -4.2] (textbook normalisation, SC_G = G/3), R in
    [0.1L, 0.35L]. The warm-start net is trained first on equilibrated fields
    of other points produced by the batched solver; the batch launch shapes
    come from the autotuner (its time is reported, not counted). train_tol /
    train_max_steps: the training sweep's tolerance and step budget (above
    4096 training points it runs on the evaluation batch's tuned launch);
    ablations=False runs the four main batched runs only."""
    from idmx import configs, sweep, tune, warmstart
    from idmx.IdpyCode import Tenet
    tenet = Tenet(0)
    Ls = tuple(range(64, Lmax + 1, 32))
    train_pts = sweep.sample_points(n_train, configs.SEED0 + 5, Ls=Ls)
    eval_pts = sweep.sample_points(n_eval, configs.SEED0 + 4, Ls=Ls)
    out = {'n_train': n_train, 'n_eval': n_eval, 'L': [Ls[0], Ls[-1]]}
    t0 = time.perf_counter()
    l64, tim64 = tune.tune_batch(tenet, eval_pts, 'fp64', use_cache=False)
    l32, tim32 = tune.tune_batch(tenet, eval_pts, 'fp32s', use_cache=False)
    s64 = tune.tune_batch_small(tenet, eval_pts, 'fp64')
    s32 = tune.tune_batch_small(tenet, eval_pts, 'fp32s')
    out['tune_s'] = time.perf_counter() - t0
    out['launch_small_fp64'] = {'best': s64[0], 'sites': s64[1]}
    out['launch_small_fp32s'] = {'best': s32[0], 'sites': s32[1]}
    out['launch_fp64'] = {'best': l64, 'ms_per_step': tim64[','.join(map(str, l64))]}
    out['launch_fp32s'] = {'best': l32, 'ms_per_step': tim32[','.join(map(str, l32))]}
    log('batch launch fp64 %s %.4f ms/step, fp32s %s %.4f ms/step' % (
        l64, out['launch_fp64']['ms_per_step'], l32, out['launch_fp32s']['ms_per_step']))
    t0 = time.perf_counter()
    if n_train > 4096:
        lt, lts = l64, s64
    else:
        lt, _ = tune.tune_batch(tenet, train_pts, 'fp64', use_cache=False)
        lts = tune.tune_batch_small(tenet, train_pts, 'fp64')
    out['train_tol'], out['train_max_steps'] = train_tol, train_max_steps
    res_tr, f_tr = sweep.run_sweep(
        tenet, train_pts, mode='fp64', tol=train_tol, max_steps=train_max_steps, launch=lt,
        launch_small=lts,
        progress=lambda lab, st, live, dn: log('training sweep: %d steps, %d live, max |dn| %.1e'
                                               % (st, live, dn)))
    out['train_data_s'] = time.perf_counter() - t0
    log('training fields: %.1f s, converged %d/%d' % (out['train_data_s'],
                                                       res_tr.converged.sum(), n_train))
    t0 = time.perf_counter()
    conv_fields = [f if c else None for f, c in zip(f_tr, res_tr.converged)]
    profile = warmstart.ProfileModel()
    out['profile_fields_used'] = profile.fit(train_pts, conv_fields)
    out['uniform_fields_used'] = profile.n_uniform
    X, Y, keep = warmstart.training_set(train_pts, conv_fields, profile)
    out['train_fields_used'] = len(keep)
    net, hist = warmstart.train(tenet, X, Y, epochs=epochs, batch=32, log=log)
    out['train_s'] = time.perf_counter() - t0
    out['train_loss'] = [float(hist[0]), float(hist[-1])]
    # plain per-point FP64 runs from the cold start
    steps_p, conv_p, wall_p, f_p = sweep.plain_runs(
        tenet, eval_pts, progress=lambda i, w: log('plain runs: %d points, %.0f s' % (i, w)))
    out['plain_fp64'] = {'wall_s': wall_p, 'steps_mean': float(steps_p.mean()),
                         'steps_max': int(steps_p.max()), 'converged': int(conv_p.sum())}
    log('plain: %.1f s' % wall_p)
    t0 = time.perf_counter()
    starts = warmstart.predict_fields(tenet, net, eval_pts, profile=profile)
    t_inf = time.perf_counter() - t0
    t0 = time.perf_counter()
    masses = [float(np.sum(sweep.cold_start(p))) for p in eval_pts]
    prof_starts = profile.predict(eval_pts, masses)
    t_prof = time.perf_counter() - t0
    out['starts_uniform'] = int(sum(profile.is_uniform(p, m) for p, m in zip(eval_pts, masses)))
    # every run uses the default fallback controller (mass screen, non-finite
    # reset, growth rule: 5 checks above 10x the best residual; a diverging
    # cold start restarts once from the gentle start) and writes learned
    # starts at rest under the Guo forcing; the last four runs are
    # ablations: the cold start without the gentle restart (the plain runs'
    # rules: same steps, same points), the profile model alone (no net), the
    # reference's f = w n populations, and the growth rule off
    runs = {'batched_cold_fp64': (None, 'fp64', 5, 'force', 0.0, True),
            'batched_cold': (None, 'fp32+fp64', 5, 'force', 0.0, True),
            'batched_warm_fp64': (starts, 'fp64', 5, 'force', t_inf, True),
            'batched_warm': (starts, 'fp32+fp64', 5, 'force', t_inf, True),
            'batched_cold_fp64_no_gentle': (None, 'fp64', 5, 'force', 0.0, False),
            'batched_warm_fp64_profile_only': (prof_starts, 'fp64', 5, 'force', t_prof, True),
            'batched_warm_fp64_plain_populations': (starts, 'fp64', 5, 'plain', t_inf, True),
            'batched_warm_fp64_no_growth_rule': (starts, 'fp64', None, 'force', t_inf, True)}
    if not ablations:
        runs = {k: v for k, v in runs.items()
                if k in ('batched_cold_fp64', 'batched_cold', 'batched_warm_fp64', 'batched_warm')}
    fields = {}
    for key, (st, mode, gc, rest, t_start, gentle) in runs.items():
        res, f = sweep.run_sweep(tenet, eval_pts, start=st, mode=mode, launch=l64,
                                 launch32=l32, launch_small=s64, launch32_small=s32,
                                 grow_checks=gc, rest_init=rest, gentle=gentle)
        fields[key] = f
        wall = res.wall_s + t_start
        out[key] = {'wall_s': wall, 'gpu_ms': res.gpu_ms, 'steps_mean': float(res.steps.mean()),
                    'steps_max': int(res.steps.max()), 'converged': int(res.converged.sum()),
                    'fallbacks': int(res.fallback.sum()),
                    'fallback_reasons': res.reset_counts(),
                    'history': res.history[::max(1, len(res.history) // 12)] + res.history[-1:]}
        if st is not None:
            out[key]['inference_s'] = t_start
        log('%s: %.1f s, steps mean %.0f, converged %d' % (key, wall, res.steps.mean(),
                                                           res.converged.sum()))
    # initial-field error of warm vs cold start against the converged fields
    err0 = lambda st, f: float(np.abs(st.ravel() - f.ravel()).max())
    ok = [i for i in range(n_eval) if conv_p[i]]
    out['start_err_cold'] = float(np.median([err0(sweep.cold_start(eval_pts[i]), f_p[i])
                                             for i in ok]))
    out['start_err_warm'] = float(np.median([err0(starts[i], f_p[i]) for i in ok]))
    # Laplace pressures against the plain runs, over the points whose end
    # state is a centred droplet in both runs. Near the droplet / stripe
    # transition (R_eq ~ L / pi) a symmetric droplet can leave the droplet
    # branch on a rounding-seeded asymmetry, so the end state there depends
    # on the arithmetic's rounding and is not unique.
    dp = lambda f, p: sweep.laplace_dp(f, p.G)
    drop_p = [sweep.is_centred_droplet(f) for f in f_p]
    out['plain_end_states_droplet'] = int(sum(drop_p))
    for key in runs:
        both = [a and sweep.is_centred_droplet(fw) for a, fw in zip(drop_p, fields[key])]
        idx = [i for i, b in enumerate(both) if b]
        rel = np.array([abs(dp(fields[key][i], eval_pts[i]) - dp(f_p[i], eval_pts[i]))
                        / abs(dp(f_p[i], eval_pts[i])) for i in idx])
        out[key]['droplet_in_both'] = int(sum(both))
        out[key]['laplace_rel_diff_max'] = float(rel.max()) if rel.size else None
        if rel.size:
            w = idx[int(rel.argmax())]
            q = eval_pts[w]
            out[key]['laplace_rel_diff_max_point'] = {
                'index': w, 'G': q.G, 'R': q.R, 'L': q.L, 'plain_steps': int(steps_p[w]),
                'centre_offset': [float(v) for v in sweep.droplet_centre(fields[key][w])],
                'plain_centre_offset': [float(v) for v in sweep.droplet_centre(f_p[w])]}
    for key in runs:
        out['speedup_%s_vs_plain' % key] = wall_p / out[key]['wall_s']
    out['speedup_warm_vs_plain'] = max(out['speedup_batched_warm_vs_plain'],
                                       out['speedup_batched_warm_fp64_vs_plain'])
    net.free()
    tenet.End()
    return out


def run_precision_study(steps=5000):
    """FP64 vs FP32-shifted populations on the D3Q19 MRT 256^3 droplet
    (sphere R = 40, SC_G = -5/3) after `steps` steps: liquid and gas
    densities (n at the centre and at site V-1, the Laplace probes), Laplace
    pressure, surface tension from Laplace's law with the equimolar radius
    (sigma = dP R / 2), maximum Mach number of the spurious currents, mass
    drift, and the largest density difference between the two runs. The FP64
    path is the one pinned to the oracle."""
    from idmx import configs
    from idmx.IdpyCode import IdpyMemory, Tenet
    from idmx.LBM import hermite
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    w = WORKLOADS['d3q19_mrt_256']
    tenet = Tenet(0)
    n0 = configs.sphere3d()
    mrt = hermite.mrt_kernel_tables(w['xi'], mrt_S(w['xi']))
    G = configs.SC_G
    psi = lambda x: 1.0 - np.exp(-x)
    press = lambda x: x / 3.0 + 0.5 * G * psi(x) ** 2
    out, fields = {}, {}
    for prec in ('fp64', 'fp32s'):
        eng = SCEngine(tenet, w['xi'], w['es'], w['dims'], 'exp', G, 1.0, precision=prec,
                       rho0=1.0 if prec == 'fp32s' else 0.0, mrt=mrt)
        V, D = eng.V, eng.D
        n = IdpyMemory.OnDevice(n0, tenet)
        ps = IdpyMemory.Zeros(V, np.float64, tenet)
        u = IdpyMemory.Zeros(V * D, np.float64, tenet)
        pop, swap = PopArray(eng, tenet), PopArray(eng, tenet)
        eng.init_pop(pop, n, u)
        m0, _ = eng.mass_momentum(pop)
        t0 = time.perf_counter()
        if eng.run(pop, swap, steps):
            pop, swap = swap, pop
        tenet.Sync()
        wall = time.perf_counter() - t0
        m1, _ = eng.mass_momentum(pop)
        eng.moments(pop, n, ps, u)
        mass = IdpyMemory.Sum(n)
        nh = n.D2H()
        u2 = (u.D2H().reshape(D, V) ** 2).sum(axis=0)
        c = 128 + 256 * (128 + 256 * 128)
        n_l, n_g = float(nh[c]), float(nh[V - 1])
        dP = press(n_l) - press(n_g)
        R_eq = (3.0 * (mass - V * n_g) / (4.0 * np.pi * (n_l - n_g))) ** (1.0 / 3.0)
        out[prec] = {'n_l': n_l, 'n_g': n_g, 'dP': dP, 'R_eq': R_eq, 'sigma': 0.5 * dP * R_eq,
                     'max_mach': float(np.sqrt(u2.max() * 3.0)),
                     'mass_rel_drift': abs(m1 - m0) / m0, 'wall_s': wall}
        fields[prec] = nh
        for x in (n, ps, u, pop, swap):
            x.Free()
        eng.close()
    a, b = out['fp64'], out['fp32s']
    out['fp32s_vs_fp64'] = {
        'max_abs_dn': float(np.abs(fields['fp64'] - fields['fp32s']).max()),
        'rel_n_l': abs(b['n_l'] - a['n_l']) / a['n_l'],
        'rel_n_g': abs(b['n_g'] - a['n_g']) / a['n_g'],
        'rel_sigma': abs(b['sigma'] - a['sigma']) / abs(a['sigma']),
        'rel_max_mach': abs(b['max_mach'] - a['max_mach']) / a['max_mach']}
    out['steps'] = steps
    tenet.End()
    return out


def run_sweep_batch_only(n_points, Lmax, mode='fp64', max_steps=400000, plain_sample=0,
                         learned=0, epochs=120):
    """The batched sweep alone at a larger scale (the per-point baseline
    would take hours there): n_points centred droplets, L uniform in
    {64, 96, ..., Lmax}, G in [-5.5, -4.2] (SC_G = G/3), R in [0.1 L, 0.35 L],
    from the cold start to max |dn| < 1e-10 between checks 200 steps apart,
    at most max_steps steps per point. Reports the step budget the points
    needed (quantiles, the largest per lattice size, the points left) and,
    with plain_sample = K, the plain per-point FP64 runs of K points spread
    over the sizes, for comparison (steps, time, Laplace pressure).
    learned = M > 0: the points start from the learned warm start, its
    profile model and net fitted on M other points of the same sizes that
    the batched solver relaxes to 1e-8 first (at most 300,000 steps each;
    not counted in the wall time, the prediction is)."""
    from idmx import configs, sweep, tune, warmstart
    from idmx.IdpyCode import Tenet
    tenet = Tenet(0)
    pts = sweep.sample_points(n_points, configs.SEED0 + 6, Ls=tuple(range(64, Lmax + 1, 32)))
    prec = 'fp64' if mode == 'fp64' else 'fp32s'
    t0 = time.perf_counter()
    l_full, tim = tune.tune_batch(tenet, pts, 'fp64', use_cache=False)
    l_small = tune.tune_batch_small(tenet, pts, 'fp64')
    kw = {'launch': l_full, 'launch_small': l_small}
    if mode != 'fp64':
        kw['launch32'], _ = tune.tune_batch(tenet, pts, 'fp32s', use_cache=False)
        kw['launch32_small'] = tune.tune_batch_small(tenet, pts, 'fp32s')
    t_tune = time.perf_counter() - t0
    log = lambda *h: print('progress', *h, file=sys.stderr, flush=True)
    starts, t_inf, info = None, 0.0, None
    if learned:
        t0 = time.perf_counter()
        tr = sweep.sample_points(learned, configs.SEED0 + 7, Ls=tuple(range(64, Lmax + 1, 32)))
        res_tr, f_tr = sweep.run_sweep(tenet, tr, mode='fp64', tol=1e-8, max_steps=300000,
                                       launch=l_full, launch_small=l_small, progress=log)
        t_data = time.perf_counter() - t0
        t0 = time.perf_counter()
        conv = [f if c else None for f, c in zip(f_tr, res_tr.converged)]
        del f_tr
        profile = warmstart.ProfileModel()
        n_prof = profile.fit(tr, conv)
        X, Y, keep = warmstart.training_set(tr, conv, profile)
        del conv
        net, hist = warmstart.train(tenet, X, Y, epochs=epochs, batch=32,
                                    log=lambda m: log(m))
        del X, Y
        t_train = time.perf_counter() - t0
        t0 = time.perf_counter()
        starts = warmstart.predict_fields(tenet, net, pts, profile=profile)
        t_inf = time.perf_counter() - t0
        net.free()
        info = {'train_points': learned, 'train_converged': int(res_tr.converged.sum()),
                'profile_fields_used': n_prof, 'net_fields_used': len(keep),
                'train_data_s': t_data, 'train_s': t_train, 'inference_s': t_inf,
                'train_loss': [float(hist[0]), float(hist[-1])],
                'profile_starts': int(sum(p is not None for p in
                                          profile.predict(pts, [float(np.sum(sweep.cold_start(q)))
                                                                for q in pts])))}
    res, fields = sweep.run_sweep(tenet, pts, start=starts, mode=mode, max_steps=max_steps,
                                  **kw, progress=log)
    out = {'n_points': n_points, 'L': [64, Lmax], 'sites': int(sum(p.L * p.L for p in pts)),
           'mode': mode, 'precision_batch': prec, 'tune_s': t_tune, 'launch': l_full,
           'ms_per_step_all_live': tim[','.join(map(str, l_full))],
           'launch_small': l_small, 'wall_s': res.wall_s + t_inf, 'gpu_ms': res.gpu_ms,
           'start': 'learned' if learned else 'cold', 'learned': info,
           'fallback_reasons': res.reset_counts(),
           'steps_mean': float(res.steps.mean()), 'steps_max': int(res.steps.max()),
           'converged': int(res.converged.sum()), 'max_steps': max_steps,
           'history': res.history[::max(1, len(res.history) // 16)] + res.history[-1:]}
    ok = res.converged
    if ok.any():
        out['steps_quantiles'] = {q: int(np.quantile(res.steps[ok], q / 100))
                                  for q in (50, 90, 99, 100)}
    out['steps_max_by_L'] = {int(L): int(max(res.steps[i] for i in range(len(pts))
                                             if pts[i].L == L and ok[i]) if any(
                                                 pts[i].L == L and ok[i] for i in range(len(pts)))
                                         else 0)
                             for L in sorted({p.L for p in pts})}
    out['not_converged'] = [{'L': pts[i].L, 'G': round(pts[i].G, 4), 'R': round(pts[i].R, 2),
                             'steps': int(res.steps[i]), 'finite': bool(np.isfinite(fields[i]).all())}
                            for i in range(len(pts)) if not ok[i]]
    if plain_sample:
        order = sorted(range(len(pts)), key=lambda i: (pts[i].L, i))
        pick = [order[int(k * (len(order) - 1) / max(1, plain_sample - 1))]
                for k in range(plain_sample)]
        sub = [pts[i] for i in pick]
        st_p, conv_p, wall_p, f_p = sweep.plain_runs(tenet, sub, max_steps=max_steps)
        rows = []
        for j, i in enumerate(pick):
            row = {'L': sub[j].L, 'steps_plain': int(st_p[j]), 'steps_batched': int(res.steps[i]),
                   'converged_plain': bool(conv_p[j]), 'converged_batched': bool(ok[i])}
            if conv_p[j] and ok[i] and sweep.is_centred_droplet(f_p[j]) \
                    and sweep.is_centred_droplet(fields[i]):
                a, b = sweep.laplace_dp(f_p[j], sub[j].G), sweep.laplace_dp(fields[i], sub[j].G)
                row['laplace_rel_diff'] = float(abs(a - b) / abs(a))
            rows.append(row)
        out['plain_sample'] = {'points': rows, 'wall_s': wall_p}
    tenet.End()
    return out


def run_batch_step_study(n_points, Lmax, steps, launch=None):
    """The batched FP64 step of the sweep study with every domain live: its
    n_points centred droplets (L uniform in {64, 96, ..., Lmax}, cold start),
    the autotuner's launch (or `launch` = (kind, tx, H, grid)), time per step
    from HIP events on the library's stream as the difference of a run of
    2 * steps and a run of `steps` (each run twice, the second timed: the
    first captures its graph; the difference cancels the pull form's
    conversions). Algorithmic bytes 144 per site (nine FP64 populations read
    once and written once)."""
    from idmx import configs, sweep, tune
    from idmx.IdpyCode import Tenet
    tenet = Tenet(0)
    pts = sweep.sample_points(n_points, configs.SEED0 + 4, Ls=tuple(range(64, Lmax + 1, 32)))
    if launch is None:
        launch, _ = tune.tune_batch(tenet, pts, 'fp64', use_cache=False)
    b = sweep.Batch(tenet, pts, 'fp64')
    for i, p in enumerate(pts):
        b.set_density(i, sweep.cold_start(p))
    b.set_launch(*launch)
    b.run(steps)
    t_a = b.run(steps)
    b.run(2 * steps)
    t_b = b.run(2 * steps)
    ms = (t_b - t_a) / steps
    sites = int(sum(p.L * p.L for p in pts))
    gbs = 144 * sites / (ms / 1e3) / 1e9
    out = {'n_points': n_points, 'L': [64, Lmax], 'sites': sites, 'launch': list(launch),
           'work_items': b.get_launch()[4], 'steps_timed': steps, 'ms_per_step': ms,
           'mlups': sites / (ms / 1e3) / 1e6, 'achieved_GBs': gbs,
           'hbm_frac': gbs / PEAK_HBM_GBS, 'bytes_per_site': 144}
    b.close()
    tenet.End()
    return out


def calibrate(reps=10):
    """Plain copy of 19 x 256^3 float64 values (2.55 GB read + 2.55 GB
    written per launch): achievable HBM bandwidth and the reference launch
    for the FETCH_SIZE / WRITE_SIZE counter scale."""
    import ctypes as C
    from idmx import _lib
    from idmx.IdpyCode import IdpyMemory, Tenet
    t = Tenet(0)
    n = 19 * 256 ** 3
    a = IdpyMemory.Zeros(n, np.float64, t)
    b = IdpyMemory.Zeros(n, np.float64, t)
    ms = C.c_double()
    best = []
    for _ in range(reps):
        _lib.call('idmx_stream_copy_f64', t.handle, b.ptr, a.ptr, n, C.byref(ms))
        best.append(ms.value)
    a.Free()
    b.Free()
    t.End()
    med = float(np.median(best))
    return {'bytes_per_launch': 16 * n, 'ms_median': med,
            'GBs': 16 * n / (med / 1e3) / 1e9}


def tune_report(names):
    """The autotuner's timing table (ms per step, % of roofline) for each
    workload."""
    from idmx import configs, tune
    from idmx.IdpyCode import Tenet
    from idmx.LBM import hermite
    from idmx.LBM._engine import SCEngine
    t = Tenet(0)
    for name in names:
        w = WORKLOADS[name]
        mrt = hermite.mrt_kernel_tables(w['xi'], mrt_S(w['xi'])) if w['coll'] == 'mrt' else None
        eng = SCEngine(t, w['xi'], w['es'], w['dims'], 'exp', configs.SC_G, 1.0,
                       precision=w['prec'], rho0=1.0 if w['prec'] == 'fp32s' else 0.0, mrt=mrt)
        best, tim = tune.tune_sim(eng, steps=5, use_cache=False)
        B = 2 * eng.Q * (8 if w['prec'] == 'fp64' else 4)
        print('%s best %s' % (name, best), flush=True)
        for k, ms in sorted(tim.items(), key=lambda kv: kv[1]):
            print('  %-16s %.4f ms  %.1f%%' % (k, ms, 100 * B * eng.V / (ms / 1e3) / 8e12),
                  flush=True)
        eng.close()
    t.End()


def traffic_from_profiles(name):
    """HBM bytes per launch of the workload's fused kernel from the committed
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json, scaled by the
    calibration copy), with the launch shape it was measured at; None if
    the workload was not profiled."""
    p = os.path.join(ROOT, 'profiles', 'traffic.json')
    try:
        t = json.load(open(p)).get(name)
    except (OSError, ValueError):
        return None, None
    if not t:
        return None, None
    return float(t['bytes_per_launch']), t.get('launch')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--workload', default='d3q19_mrt_256', choices=sorted(WORKLOADS))
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--extra', default='d2q9_4096_e8,d3q19_bgk_256,d3q19_mrt_256_fp32s',
                    help='comma-separated extra workloads reported under "extra"')
    ap.add_argument('--no-tune', action='store_true',
                    help='skip the launch-shape autotuner (use the creation default)')
    ap.add_argument('--launch', default=None,
                    help='fixed launch shape tx,ty,H,hist for the main workload')
    ap.add_argument('--calibrate', action='store_true',
                    help='time the plain copy kernel (bandwidth / counter calibration)')
    ap.add_argument('--tune-report', default=None,
                    help='print the autotuner table for these comma-separated workloads')
    ap.add_argument('--precision-study', type=int, default=0, metavar='STEPS',
                    help='FP64 vs FP32-shifted observables on the D3Q19 MRT 256^3 droplet '
                         'after STEPS steps')
    ap.add_argument('--two-pass-study', action='store_true',
                    help='time the two-kernel step (the reference kernel structure) on each '
                         'workload instead of the step benchmark')
    ap.add_argument('--sweep', action='store_true',
                    help='run the (G, R, L) sweep study (time to steady state, learned '
                         'warm start vs plain) instead of the step benchmark')
    ap.add_argument('--sweep-batch-only', type=int, default=0, metavar='N',
                    help='time the batched sweep alone on N points (--sweep-Lmax, '
                         '--sweep-mode)')
    ap.add_argument('--batch-step-study', type=int, default=0, metavar='STEPS',
                    help='time the batched FP64 step with every domain of the sweep '
                         'study live (--sweep-eval points, --sweep-Lmax) over STEPS steps')
    ap.add_argument('--batch-launch', default=None,
                    help='kind,tx,H,grid for --batch-step-study (default: autotuned)')
    ap.add_argument('--sweep-mode', default='fp64', choices=['fp64', 'fp32+fp64'])
    ap.add_argument('--sweep-max-steps', type=int, default=400000,
                    help='step budget per point of --sweep-batch-only')
    ap.add_argument('--sweep-plain-sample', type=int, default=0, metavar='K',
                    help='with --sweep-batch-only: plain per-point runs of K points')
    ap.add_argument('--sweep-learned', type=int, default=0, metavar='M',
                    help='with --sweep-batch-only: start from the learned warm start, '
                         'fitted on M other points of the same sizes')
    ap.add_argument('--sweep-train', type=int, default=512)
    ap.add_argument('--sweep-train-tol', type=float, default=1e-8,
                    help='tolerance of the training sweep of --sweep')
    ap.add_argument('--sweep-train-max-steps', type=int, default=400000,
                    help='step budget per training point of --sweep')
    ap.add_argument('--no-sweep-ablations', action='store_true',
                    help='--sweep: the four main batched runs only')
    ap.add_argument('--sweep-eval', type=int, default=128)
    ap.add_argument('--sweep-Lmax', type=int, default=256)
    ap.add_argument('--sweep-epochs', type=int, default=120)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: the library has no multi-GPU path '
              '(a lattice would be cut into slabs along its slowest axis, one per GPU); '
              'run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get('RANK', '0'))
    if args.calibrate:
        print(json.dumps({'calibration': calibrate()}), flush=True)
        return
    if args.precision_study:
        print(json.dumps({'precision_study': run_precision_study(args.precision_study)}),
              flush=True)
        return
    if args.tune_report:
        tune_report(args.tune_report.split(','))
        return
    if args.two_pass_study:
        # the reference's kernel structure (density/psi pass, then force-
        # collide-stream) on the GPU: 3 Q s + 32 bytes per site (n and psi
        # written, populations read twice, n and psi read back), against the
        # fused step's 2 Q s
        out = {}
        for x in [args.workload] + [e for e in args.extra.split(',') if e and e != args.workload]:
            rx = run_gpu(x, 50, 5, tune=False, two_pass=True)
            s = rx['B_alg'] // (2 * rx['Q'])
            ref_b = 3 * rx['Q'] * s + 32
            out[x] = {'mlups': round(rx['mlups'], 1), 'kernel_ms': round(rx['kernel_ms'], 4),
                      'frac_of_fused_bytes': round(rx['B_alg'] * rx['V'] / (rx['kernel_ms'] / 1e3)
                                                   / 1e9 / PEAK_HBM_GBS, 4),
                      'structure_bytes_per_site': ref_b,
                      'frac_of_structure_bytes': round(ref_b * rx['V'] / (rx['kernel_ms'] / 1e3)
                                                       / 1e9 / PEAK_HBM_GBS, 4)}
        print(json.dumps({'two_pass_study': out}), flush=True)
        return
    if args.batch_step_study:
        bl = tuple(int(v) for v in args.batch_launch.split(',')) if args.batch_launch else None
        r = run_batch_step_study(args.sweep_eval, args.sweep_Lmax, args.batch_step_study, bl)
        print(json.dumps({'batch_step_study': r}), flush=True)
        return
    if args.sweep_batch_only:
        r = run_sweep_batch_only(args.sweep_batch_only, args.sweep_Lmax, args.sweep_mode,
                                 args.sweep_max_steps, args.sweep_plain_sample,
                                 args.sweep_learned, args.sweep_epochs)
        print(json.dumps({'sweep_batch_only': r}), flush=True)
        return
    if args.sweep:
        r = run_sweep_study(args.sweep_train, args.sweep_eval, args.sweep_Lmax,
                            args.sweep_epochs, log=lambda m: print(m, file=sys.stderr, flush=True),
                            train_tol=args.sweep_train_tol,
                            train_max_steps=args.sweep_train_max_steps,
                            ablations=not args.no_sweep_ablations)
        line = {'metric': 'sweep time-to-steady-state learned vs plain',
                'value': round(r['speedup_warm_vs_plain'], 3), 'unit': 'x (plain / learned)',
                'higher_is_better': True, 'n_gpus': 1, 'data': 'synthetic',
                'dtype': 'f32+f64', 'config': {'workload': 'd2q9_sweep'}, 'study': r}
        print(json.dumps(line), flush=True)
        return
    name = args.workload
    w = WORKLOADS[name]
    launch = tuple(int(v) for v in args.launch.split(',')) if args.launch else None
    r = run_gpu(name, args.steps, args.warmup, tune=not args.no_tune, launch=launch)
    achieved = r['B_alg'] * r['V'] / (r['kernel_ms'] / 1e3) / 1e9
    traffic, traffic_launch = traffic_from_profiles(name)
    line = {
        'metric': 'MLUPS & %HBM roofline (D3Q19 SC)',
        'value': round(r['mlups'], 1),
        'unit': 'MLUPS',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(r['wall_s'] * 1e3 / args.steps, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if w['prec'] == 'fp64' else 'f32',
        'data': 'synthetic',
        'config': workload_desc(name, w),
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4),
                     'traffic': traffic, 'traffic_launch': traffic_launch,
                     'bytes_per_site': r['B_alg'], 'kernel_ms': round(r['kernel_ms'], 4)},
        'mass_rel_drift': r['mass_rel_drift'],
        'launch': r['launch'],
        'device': r['device'],
        'device_clocks': device_clocks(),
    }
    extra = {}
    for x in [e for e in args.extra.split(',') if e and e != name]:
        rx = run_gpu(x, max(20, args.steps // 4), max(5, args.warmup // 4),
                     tune=not args.no_tune)
        ax = rx['B_alg'] * rx['V'] / (rx['kernel_ms'] / 1e3) / 1e9
        tx, tlx = traffic_from_profiles(x)
        extra[x] = {'mlups': round(rx['mlups'], 1), 'hbm_frac': round(ax / PEAK_HBM_GBS, 4),
                    'achieved_GBs': round(ax, 1), 'kernel_ms': round(rx['kernel_ms'], 4),
                    'mass_rel_drift': rx['mass_rel_drift'], 'traffic': tx,
                    'traffic_launch': tlx,
                    'launch': [rx['launch'][k] for k in ('tile_x', 'tile_y',
                                                         'planes_per_block', 'variant')]}
    if extra:
        line['extra'] = extra
    if rank == 0 and not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(name)
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
