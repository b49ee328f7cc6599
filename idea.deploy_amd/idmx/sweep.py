"""Batched (G, R, L) parameter sweep of centred Shan-Chen droplets.

The reference runs one ShanChenMultiPhase per parameter point and stops it
with a convergence callback between checkpoints
(idpy/LBM/MultiPhase.py:1015-1097, MultiPhaseLoopChecks.py:96-163). Here all
points of a sweep live in one idmx batch (libidmx.so, idmx_batch_*): one
launch per step advances every live domain, a captured graph replays 100
steps, and every `check_every` steps a device reduction returns each domain's
max |n - n_prev| (FP64). A domain is steady when that falls below `tol`; it
is then dropped from the work list, so the remaining domains get the GPU.

Modes
  cold FP64:   tanh droplet start, FP64 populations to tol
  FP32 + FP64: any start, FP32 shifted populations to tol32, then the state
               is converted to FP64 and relaxed to tol (FP64 fixes the final
               accuracy)
Fallback controller (for caller-supplied starts; a cold start is never reset,
its trajectory would repeat):
  - a start whose total mass differs from the cold start's by more than
    `mass_tol` (relative) describes another droplet and is replaced by the
    cold start before the run; within it, the start is shifted uniformly to
    the cold start's mass (the dynamics conserve mass, so that fixes the
    steady state);
  - a start that turns non-finite is reset to the cold start;
  - a start whose residual grows is reset to the cold start: `grow_checks`
    consecutive checks with max |dn| above ten times its best value so far
    and above its value at the first check after the start. The second
    condition keeps the transient rise of a good start (the relaxation is not
    monotone) from triggering it.
A reset domain gets a fresh step budget, so resets cost time but not
converged points. SweepResult.reason records why each point fell back.
A cold start that turns non-finite (near the weak end of the G range its
sharp interface drives populations negative) is restarted once from
gentle_start, the same mass spread almost uniformly (run_sweep's `gentle`;
SweepResult.gentle marks those points); it is dropped as failed only if that
diverges too.
"""
import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .configs import droplet

PREC = {'fp64': 0, 'fp32s': 1}
E_IDS = {'E4': 4, 'E6': 6, 'E8': 8}


@dataclass
class SweepPoint:
    L: int
    G: float          # SC_G in the FStencils (e2 = 1) normalisation
    R: float
    tau: float = 1.0


# SweepResult.reason codes
FB_NONE, FB_NONFINITE, FB_GROWTH, FB_MASS = 0, 1, 2, 3


@dataclass
class SweepResult:
    steps: np.ndarray
    converged: np.ndarray
    fallback: np.ndarray
    wall_s: float
    gpu_ms: float
    history: list = field(default_factory=list)
    reason: np.ndarray = None   # FB_* per point
    gentle: np.ndarray = None   # restarted from gentle_start after the cold start diverged

    def reset_counts(self):
        r = self.reason if self.reason is not None else np.zeros(len(self.steps), dtype=int)
        g = self.gentle if self.gentle is not None else np.zeros(len(self.steps), dtype=bool)
        return {'nonfinite': int((r == FB_NONFINITE).sum()), 'growth': int((r == FB_GROWTH).sum()),
                'mass': int((r == FB_MASS).sum()), 'gentle': int(g.sum())}


def cold_start(p, W=2.0):
    return droplet((p.L, p.L), p.R, W)


def gentle_start(p, amp=0.05, W=2.0):
    """The cold start's mass spread almost uniformly: its mean density plus
    a centred bump of the cold droplet's shape, peak-to-peak `amp`. Near the
    weak end of the G range the cold start's sharp 0.1 / 2.0 interface drives
    the populations negative within ~2000 steps, although the steady state
    there is often a uniform gas; from this start the dynamics stay small."""
    n0 = cold_start(p, W)
    shape = (n0 - n0.min()) / (n0.max() - n0.min())
    shape = shape - shape.mean()
    return n0.mean() + amp * shape / (shape.max() - shape.min())


class Batch:
    """ctypes handle of one idmx_batch_t."""

    def __init__(self, tenet, points, precision='fp64', estencil='E4', rho0=1.0):
        self.tenet, self.points = tenet, list(points)
        arr = (_lib.BatchDomain * len(self.points))()
        for i, p in enumerate(self.points):
            arr[i].L0 = arr[i].L1 = int(p.L)
            arr[i].SC_G = float(p.G)
            arr[i].tau = float(p.tau)
            arr[i].rho0 = float(rho0)
        h = C.c_void_p()
        _lib.call('idmx_batch_create', tenet.handle, arr, len(self.points), PREC[precision],
                  E_IDS[estencil], 0, C.byref(h))
        self.handle = h
        tenet._live[id(self)] = self
        self.n = len(self.points)

    def close(self):
        if self.handle:
            if self.tenet.handle:
                _lib.call('idmx_batch_destroy', self.handle)
            self.tenet._live.pop(id(self), None)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, kind, tx, H, grid=0):
        """kind 0: marching items of tx x H rows; kind 1: tx x H tiles;
        kinds 2 / 3: 0 / 1 on the pull form; kind 4: tiles advanced two
        steps per launch (E4); kind 5: kind 4 on the x-shifted storage
        between the launches of a replayed graph (line-aligned stores)."""
        _lib.call('idmx_batch_set_launch', self.handle, int(kind), int(tx), int(H), int(grid))

    def get_launch(self):
        """(kind, tx, H, grid, live work items)."""
        v = [C.c_int() for _ in range(5)]
        _lib.call('idmx_batch_get_launch', self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_active(self, active):
        a = np.ascontiguousarray(np.asarray(active, dtype=np.int32).ravel())
        if a.size != self.n:
            raise ValueError('set_active: %d flags for %d domains' % (a.size, self.n))
        _lib.call('idmx_batch_set_active', self.handle, a.ctypes.data_as(C.POINTER(C.c_int)))

    def set_density(self, i, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.float64).ravel())
        L = self.points[i].L
        if a.size != L * L:
            raise ValueError('set_density: %d values for a %dx%d domain' % (a.size, L, L))
        _lib.call('idmx_batch_set_density', self.handle, int(i), a.ctypes.data)

    def get_density(self, i):
        L = self.points[i].L
        out = np.empty(L * L)
        _lib.call('idmx_batch_get_density', self.handle, int(i), out.ctypes.data)
        return out.reshape(L, L)

    def run(self, nsteps):
        _lib.call('idmx_batch_run', self.handle, int(nsteps))
        ms = C.c_double()
        _lib.call('idmx_batch_last_run_ms', self.handle, C.byref(ms))
        return ms.value

    def check(self):
        out = np.empty(self.n)
        _lib.call('idmx_batch_check', self.handle, out.ctypes.data)
        return out

    def symmetrize(self, dom=-1, mass=None):
        """Projects domain dom's density (-1: every live domain) onto the
        fields symmetric about its centre, then (mass: one total per domain)
        shifts it uniformly to that mass; populations to rest equilibrium."""
        m = None
        if mass is not None:
            m = np.ascontiguousarray(np.asarray(mass, dtype=np.float64))
            if m.size != self.n:
                raise ValueError('symmetrize: one mass per domain')
        _lib.call('idmx_batch_symmetrize', self.handle, int(dom),
                  None if m is None else m.ctypes.data)

    def force_rest(self, dom=-1):
        """Populations of domain dom (-1: every live domain) set to the fluid
        at rest under the Guo forcing, f = f_eq(n, -F/(2n)), from its
        check-reference density (idmx_batch_force_rest)."""
        _lib.call('idmx_batch_force_rest', self.handle, int(dom))

    def get_populations(self, i):
        """float64 (9, L*L) populations of domain i (shift removed)."""
        L = self.points[i].L
        out = np.empty((9, L * L))
        _lib.call('idmx_batch_get_populations', self.handle, int(i), out.ctypes.data)
        return out

    def copy_state_from(self, other, rest=False):
        """Populations of `other` (converted); rest=True writes the rest
        equilibrium of its density instead (zero momentum)."""
        _lib.call('idmx_batch_copy_state', self.handle, other.handle, 1 if rest else 0)


def _apply_launch(batch, active, launch, launch_small):
    """launch (the batch's tuned shape) while the live domains hold more
    sites than launch_small = (shape, sites) was tuned for, then the small
    shape: a launch tuned for a full, bandwidth-bound batch is latency-bound
    once only a few domains remain."""
    if launch_small is None:
        return
    live = sum(batch.points[i].L ** 2 for i in np.nonzero(active)[0])
    want = launch_small[0] if live <= launch_small[1] else launch
    if want is not None and tuple(want) != batch.get_launch()[:4]:
        batch.set_launch(*want)


def _relax(batch, active, warm, steps, tol, check_every, max_steps, grow_checks, hist,
           restart, label, plateau=False, launch=None, launch_small=None, plateau_checks=5,
           progress=None, reason=None, restart_gentle=None, gentled=None):
    """Runs the live domains of `batch` until each one's max |dn| < tol, or
    until every live domain has run max_steps steps since its (re)start in
    this call. A domain on a caller-supplied start (warm[i]) that turns
    non-finite, or (when grow_checks is not None) whose residual stays above
    ten times its best value and above its value at the first check for
    `grow_checks` consecutive checks, is handed to restart(i) (the cold-start
    fallback) once, with a fresh step budget; reason[i] (if given) records
    FB_NONFINITE or FB_GROWTH. A cold domain that turns non-finite is
    dropped as failed, unless restart_gentle is given: then it is handed to
    restart_gentle(i) once (gentled[i] records it), with a fresh step
    budget, and fails only if it turns non-finite again. With plateau=True
    (the FP32 phase) a domain also
    finishes once its residual has not improved by 20 % for `plateau_checks`
    checks: FP32 rounding sets a floor below which it cannot go, and the
    FP64 phase takes over there.
    batch needs set_active, run(nsteps) -> GPU ms and check() -> max |dn|
    per domain (NaN or inf for a non-finite state). launch / launch_small:
    see _apply_launch. progress(label, steps, live, max |dn|) is called
    every 50 checks, and after any check 30 s or more after the last call.
    Returns (still active, failed, fell back, GPU ms)."""
    n = batch.n
    if restart_gentle is not None and gentled is None:
        gentled = np.zeros(n, dtype=bool)
    best = np.full(n, np.inf)
    first = np.full(n, np.nan)   # residual at the first check after the start
    stale = np.zeros(n, dtype=int)
    grows = np.zeros(n, dtype=int)
    fell_back = np.zeros(n, dtype=bool)
    failed = np.zeros(n, dtype=bool)
    ran = np.zeros(n, dtype=np.int64)
    gpu_ms = 0.0
    t_prog = time.perf_counter()
    active = active.copy()
    batch.set_active(active)
    _apply_launch(batch, active, launch, launch_small)
    batch.check()  # reference densities
    while active.any():
        gpu_ms += batch.run(check_every)
        steps[active] += check_every
        ran[active] += check_every
        dn = batch.check()
        fin = np.isfinite(dn)
        first = np.where(active & fin & np.isnan(first), dn, first)
        grows = np.where(active & fin & (dn > 10 * best) & (dn > first) & (dn > tol),
                         grows + 1, 0)
        stale = np.where(active & fin & (dn > 0.8 * best), stale + 1, 0)
        best = np.where(active & fin, np.minimum(best, dn), best)
        done = active & fin & (dn < tol)
        if plateau:
            done |= active & fin & (stale >= plateau_checks)
        hist.append((label, int(steps[active].max()), int(active.sum()),
                     float(np.max(np.where(active & fin, dn, 0.0)))))
        if progress is not None and (len(hist) % 50 == 0
                                     or time.perf_counter() - t_prog > 30.0):
            t_prog = time.perf_counter()
            progress(*hist[-1])
        grew = grows >= grow_checks if grow_checks is not None else np.zeros(n, dtype=bool)
        bad = active & ~done & (~fin | (warm & grew))
        dead = bad & ~warm
        if restart_gentle is not None:
            soft = dead & ~gentled
            for i in np.nonzero(soft)[0]:
                restart_gentle(i)
                gentled[i] = True
                grows[i] = stale[i] = 0
                best[i] = np.inf
                first[i] = np.nan
                ran[i] = 0
            dead = dead & ~soft
        failed |= dead
        changed = (done | dead).any()
        active = active & ~done & ~dead
        for i in np.nonzero(bad & warm)[0]:
            if reason is not None:
                reason[i] = FB_NONFINITE if not fin[i] else FB_GROWTH
            restart(i)
            warm[i] = False
            fell_back[i] = True
            grows[i] = stale[i] = 0
            best[i] = np.inf
            first[i] = np.nan
            ran[i] = 0
        if not active.any() or (ran[active] >= max_steps).all():
            break
        if changed:
            batch.set_active(active)
            _apply_launch(batch, active, launch, launch_small)
    return active, failed, fell_back, gpu_ms


def screen_starts(points, start, mass_tol=0.05):
    """The initial densities of a sweep: the cold start where start is None
    (or start[i] is None), else start[i], unless it is not finite or its
    mass differs from the cold start's by more than mass_tol (relative):
    then the cold start, recorded as a fallback. Returns (densities,
    warm flags, fell-back flags, FB_* reasons, cold-start masses)."""
    n = len(points)
    warm = np.array([start is not None and start[i] is not None for i in range(n)], dtype=bool)
    init = [cold_start(p) if not warm[i] else np.asarray(start[i], dtype=np.float64).ravel()
            for i, p in enumerate(points)]
    for i, p in enumerate(points):
        if init[i].size != p.L * p.L:
            raise ValueError('run_sweep: start %d has %d values for a %dx%d domain'
                             % (i, init[i].size, p.L, p.L))
    fb = np.zeros(n, dtype=bool)
    reason = np.zeros(n, dtype=int)
    mass0 = np.array([float(np.sum(cold_start(p))) for p in points])
    for i, p in enumerate(points):
        if not warm[i]:
            continue
        if not np.isfinite(init[i]).all():
            bad = FB_NONFINITE
        elif abs(float(np.sum(init[i])) - mass0[i]) > mass_tol * mass0[i]:
            bad = FB_MASS
        else:
            continue
        init[i] = cold_start(p)
        warm[i] = False
        fb[i] = True
        reason[i] = bad
    return init, warm, fb, reason, mass0


def run_sweep(tenet, points, start=None, mode='fp64', tol=1e-10, tol32=1e-6,
              check_every=200, max_steps=400000, grow_checks=5, estencil='E4',
              launch=None, launch32=None, rho0=1.0, symmetric=True, launch_small=None,
              launch32_small=None, progress=None, mass_tol=0.05, rest_init='force',
              gentle=True):
    """Relax every point to steady state; returns SweepResult and the final
    FP64 densities. start: None (cold) or a list of initial densities (None
    entries start cold). mode: 'fp64' (FP64 only) or 'fp32+fp64'.
    max_steps bounds each phase. launch / launch32: (kind, tx, H, grid) of
    the FP64 / FP32 batch (None keeps the library default); launch_small /
    launch32_small: ((kind, tx, H, grid), sites), the shape used once the
    live domains hold at most that many sites (tune.tune_batch_small).
    symmetric: the points are centred droplets, whose steady state is
    symmetric about the domain centre; warm starts and the FP32 -> FP64
    hand-over are projected onto symmetric fields (a broken symmetry makes
    the droplet drift towards a symmetric lattice position for ~1e5 steps).
    grow_checks / mass_tol: the fallback rules of the module docstring
    (grow_checks=None switches the growth rule off).
    rest_init: populations written from a density, i.e. of a caller-supplied
    start and at the FP32 -> FP64 hand-over: 'force' (default) the fluid at
    rest under the Guo forcing, f = f_eq(n, -F/(2n)) (Batch.force_rest);
    'plain' f = w n, as the cold start and the reference's InitPopulations
    at u = 0, which start every interface at the velocity F/(2n).
    gentle: a domain whose cold start turns non-finite is restarted once from
    gentle_start (same mass) instead of being dropped; SweepResult.gentle
    marks those points."""
    if rest_init not in ('force', 'plain'):
        raise ValueError("rest_init is 'force' or 'plain'")
    n = len(points)
    steps = np.zeros(n, dtype=np.int64)
    hist = []
    fail_all = np.zeros(n, dtype=bool)
    init, warm, fb_all, reason, mass0 = screen_starts(points, start, mass_tol)
    t0 = time.perf_counter()
    b64 = Batch(tenet, points, 'fp64', estencil)
    if launch:
        b64.set_launch(*launch)
    gpu_ms = 0.0

    gentled = np.zeros(n, dtype=bool)

    def restart64(i):
        n0 = cold_start(points[i])
        mass0[i] = float(np.sum(n0))
        b64.set_density(i, n0)

    def gentle64(i):
        n0 = gentle_start(points[i])
        mass0[i] = float(np.sum(n0))
        b64.set_density(i, n0)

    if mode == 'fp32+fp64':
        b32 = Batch(tenet, points, 'fp32s', estencil, rho0=rho0)
        if launch32:
            b32.set_launch(*launch32)
        for i in range(n):
            b32.set_density(i, init[i])
            if warm[i]:
                _shift_to_mass(b32, i, mass0, symmetric)
                if rest_init == 'force':
                    b32.force_rest(i)

        def restart32(i):
            n0 = cold_start(points[i])
            mass0[i] = float(np.sum(n0))
            b32.set_density(i, n0)

        def gentle32(i):
            n0 = gentle_start(points[i])
            mass0[i] = float(np.sum(n0))
            b32.set_density(i, n0)

        _, failed, fb, ms = _relax(b32, np.ones(n, dtype=bool), warm, steps, tol32, check_every,
                                   max_steps, grow_checks, hist, restart32, 'fp32',
                                   plateau=True, launch=launch32, launch_small=launch32_small,
                                   progress=progress, reason=reason,
                                   restart_gentle=gentle32 if gentle else None, gentled=gentled)
        gpu_ms += ms
        fb_all |= fb
        # FP32 rounding leaves each domain a small net momentum that FP64
        # would conserve (a droplet that never stops): hand over the density
        # only, at rest
        b64.copy_state_from(b32, rest=True)
        b32.close()
        # FP32 storage drifts the mass (~1e-5 relative over a few thousand
        # steps), which would move the steady state: restore it
        if symmetric:
            b64.symmetrize(-1, mass=mass0)
        if rest_init == 'force':
            b64.force_rest(-1)
        for i in np.nonzero(failed)[0]:   # FP32 blew up: FP64 from the cold start
            (gentle64 if gentled[i] else restart64)(i)
            fb_all[i] = True
    elif mode == 'fp64':
        for i in range(n):
            b64.set_density(i, init[i])
            if warm[i]:
                _shift_to_mass(b64, i, mass0, symmetric)
                if rest_init == 'force':
                    b64.force_rest(i)
    else:
        raise ValueError('mode is fp64 or fp32+fp64')
    rem, failed, fb, ms = _relax(b64, np.ones(n, dtype=bool), warm, steps, tol, check_every,
                                 max_steps, grow_checks, hist, restart64, 'fp64',
                                 launch=launch, launch_small=launch_small, progress=progress,
                                 reason=reason, restart_gentle=gentle64 if gentle else None,
                                 gentled=gentled)
    gpu_ms += ms
    fb_all |= fb
    fail_all |= failed
    fields = [b64.get_density(i) for i in range(n)]
    b64.close()
    wall = time.perf_counter() - t0
    return SweepResult(steps=steps, converged=~rem & ~fail_all, fallback=fb_all, wall_s=wall,
                       gpu_ms=gpu_ms, history=hist, reason=reason, gentle=gentled), fields


def _shift_to_mass(batch, i, mass0, symmetric):
    """Sets domain i's density to the cold start's mass by a uniform shift
    (with symmetric, after projecting it onto the centre-symmetric fields)."""
    if symmetric:
        batch.symmetrize(i, mass=mass0)
    else:
        n = batch.get_density(i)
        batch.set_density(i, n + (mass0[i] - n.sum()) / n.size)


def is_centred_droplet(n, tol=2.0):
    """True when an L x L density field is a disc of liquid centred on
    (L/2, L/2): liquid at the centre, gas at the corner, and the liquid
    chord through the centre along x, y and both diagonals within `tol`
    lattice units of the equimolar diameter."""
    n = np.asarray(n).reshape(int(round(np.sqrt(np.size(n)))), -1)
    L = n.shape[0]
    c = L // 2
    n_l, n_g = n[c, c], n[-1, -1]
    if not n_l > n_g:
        return False
    mid = 0.5 * (n_l + n_g)
    R = np.sqrt(max(n.sum() - n.size * n_g, 0.0) / (np.pi * (n_l - n_g)))
    k = np.arange(L)
    chords = [(n[c, :] > mid).sum(), (n[:, c] > mid).sum(),
              np.sqrt(2) * (n[k, k] > mid).sum(), np.sqrt(2) * (n[k, L - 1 - k] > mid).sum()]
    return all(abs(ch - 2 * R) <= 2 * tol for ch in chords)


def laplace_dp(n, G, psi=lambda x: 1 - np.exp(-x), c2=1 / 3, e2=1.0):
    """p(centre) - p(corner) with p = n c2 + G e2 psi(n)^2 / 2
    (idpy/LBM/MultiPhase.py:1690-1768)."""
    L = n.shape[0]
    p = lambda x: x * c2 + 0.5 * G * e2 * psi(x) ** 2
    return p(n[L // 2, L // 2]) - p(n[-1, -1])


def sample_points(count, seed, Ls=tuple(range(64, 513, 32)), G_lit=(-5.5, -4.2),
                  Rfrac=(0.1, 0.35)):
    """Sweep points: L uniform in Ls, G uniform in G_lit (textbook
    normalisation; SC_G = G * c2 for the e2 = 1 stencils), R uniform in
    [0.1 L, 0.35 L]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.choice(Ls))
        G = float(rng.uniform(*G_lit)) / 3.0
        R = float(rng.uniform(Rfrac[0] * L, Rfrac[1] * L))
        out.append(SweepPoint(L=L, G=G, R=R))
    return out


def plain_runs(tenet, points, tol=1e-10, check_every=200, max_steps=400000, starts=None,
               progress=None):
    """The unbatched baseline: one FP64 simulation per point, run the way a
    notebook runs MainLoop with a convergence callback, i.e. checkpoints of
    `check_every` steps and a device max |n - n_prev| check after each. A
    run whose residual turns non-finite stops there (as the batched sweep
    drops such a domain), so both spend the same steps on diverging points.
    progress(points done, seconds) is called after every 32 points.
    Returns (steps, converged, wall seconds, final densities)."""
    import time as _t
    from .IdpyCode import IdpyMemory
    from .LBM._base import PopArray
    from .LBM._engine import SCEngine
    steps = np.zeros(len(points), dtype=np.int64)
    conv = np.zeros(len(points), dtype=bool)
    fields = []
    t0 = _t.perf_counter()
    for i, p in enumerate(points):
        eng = SCEngine(tenet, 'D2Q9', 'D2E4', (p.L, p.L), 'exp', p.G, p.tau)
        V = p.L * p.L
        n0 = cold_start(p) if starts is None else starts[i]
        n = IdpyMemory.OnDevice(n0, tenet)
        nprev = IdpyMemory.OnDevice(n0, tenet)
        psi = IdpyMemory.Zeros(V, np.float64, tenet)
        u = IdpyMemory.Zeros(2 * V, np.float64, tenet)
        a, b = PopArray(eng, tenet), PopArray(eng, tenet)
        eng.init_pop(a, n, u)
        while steps[i] < max_steps:
            if eng.run(a, b, check_every):
                a, b = b, a
            steps[i] += check_every
            eng.density_psi(n, psi, a)
            d = IdpyMemory.MaxAbsDiff(n, nprev)
            n, nprev = nprev, n
            if d < tol:
                conv[i] = True
                break
            if not np.isfinite(d):
                break
        fields.append(nprev.D2H().reshape(p.L, p.L))
        for x in (n, nprev, psi, u, a, b):
            x.Free()
        eng.close()
        if progress is not None and (i + 1) % 32 == 0:
            progress(i + 1, _t.perf_counter() - t0)
    return steps, conv, _t.perf_counter() - t0, fields


def droplet_centre(n):
    """Offset (x, y) of a droplet's centre from the domain centre (L/2, L/2):
    the circular mean of the excess density over each axis (periodic)."""
    n = np.asarray(n, dtype=np.float64)
    L = int(round(np.sqrt(n.size)))
    n = n.reshape(L, L)
    w = n - n.min()
    ang = np.exp(2j * np.pi * np.arange(L) / L)
    out = []
    for prof in (w.sum(axis=0), w.sum(axis=1)):
        a = (np.angle((prof * ang).sum()) % (2 * np.pi)) * L / (2 * np.pi)
        out.append((a - L // 2 + L / 2) % L - L / 2)
    return tuple(out)
