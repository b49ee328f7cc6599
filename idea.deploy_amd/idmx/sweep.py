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
A start whose residual grows for `grow_checks` consecutive checks (or turns
non-finite) is reset to the cold start (the fallback of a bad warm start).
"""
import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .configs import N_G, N_L, droplet

PREC = {'fp64': 0, 'fp32s': 1}
E_IDS = {'E4': 4, 'E6': 6, 'E8': 8}


@dataclass
class SweepPoint:
    L: int
    G: float          # SC_G in the FStencils (e2 = 1) normalisation
    R: float
    tau: float = 1.0


@dataclass
class SweepResult:
    steps: np.ndarray
    converged: np.ndarray
    fallback: np.ndarray
    wall_s: float
    gpu_ms: float
    history: list = field(default_factory=list)


def cold_start(p, W=2.0):
    return droplet((p.L, p.L), p.R, W)


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
        self.n = len(self.points)

    def close(self):
        if self.handle:
            _lib.call('idmx_batch_destroy', self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, tx, H, grid=0):
        _lib.call('idmx_batch_set_launch', self.handle, int(tx), int(H), int(grid))

    def get_launch(self):
        v = [C.c_int() for _ in range(4)]
        _lib.call('idmx_batch_get_launch', self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_active(self, active):
        a = np.ascontiguousarray(np.asarray(active, dtype=np.int32))
        _lib.call('idmx_batch_set_active', self.handle, a.ctypes.data_as(C.POINTER(C.c_int)))

    def set_density(self, i, n):
        a = np.ascontiguousarray(np.asarray(n, dtype=np.float64).ravel())
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

    def copy_state_from(self, other):
        _lib.call('idmx_batch_copy_state', self.handle, other.handle)


def _relax(batch, active, steps, tol, check_every, max_steps, grow_checks, hist, restart,
           label, plateau=False):
    """Runs the live domains of `batch` until each one's max |dn| < tol.
    A domain whose residual stays above twice its best value for
    `grow_checks` consecutive checks, or turns non-finite, is handed to
    restart(i) (the cold-start fallback). With plateau=True (the FP32 phase)
    a domain also finishes once its residual is below 100 tol and has not
    improved by 20 % for `grow_checks` checks: FP32 rounding sets a floor."""
    n = batch.n
    best = np.full(n, np.inf)
    stale = np.zeros(n, dtype=int)
    grows = np.zeros(n, dtype=int)
    fell_back = np.zeros(n, dtype=bool)
    gpu_ms = 0.0
    batch.set_active(active)
    batch.check()  # reference densities
    while active.any():
        gpu_ms += batch.run(check_every)
        steps[active] += check_every
        dn = batch.check()
        fin = np.isfinite(dn)
        grows = np.where(active & fin & (dn > 2 * best) & (dn > tol), grows + 1, 0)
        stale = np.where(active & fin & (dn > 0.8 * best), stale + 1, 0)
        best = np.where(active & fin, np.minimum(best, dn), best)
        done = active & fin & (dn < tol)
        if plateau:
            done |= active & fin & (dn < 100 * tol) & (stale >= grow_checks)
        hist.append((label, int(steps.max()), int(active.sum()),
                     float(np.nanmax(np.where(active, dn, 0.0)))))
        changed = done.any()
        active = active & ~done
        for i in np.nonzero(active & ((grows >= grow_checks) | ~fin))[0]:
            restart(i)
            fell_back[i] = True
            grows[i] = stale[i] = 0
            best[i] = np.inf
        if not active.any() or (steps[active] >= max_steps).all():
            break
        if changed:
            batch.set_active(active)
    return active, gpu_ms, fell_back


def run_sweep(tenet, points, start=None, mode='fp64', tol=1e-10, tol32=1e-6,
              check_every=200, max_steps=400000, grow_checks=5, estencil='E4',
              launch=None, launch32=None, rho0=1.0):
    """Relax every point to steady state; returns SweepResult and the final
    FP64 densities. start: None (cold) or a list of initial densities.
    mode: 'fp64' (FP64 only) or 'fp32+fp64'."""
    n = len(points)
    steps = np.zeros(n, dtype=np.int64)
    hist = []
    fb_all = np.zeros(n, dtype=bool)
    t0 = time.perf_counter()
    b64 = Batch(tenet, points, 'fp64', estencil)
    if launch:
        b64.set_launch(*launch)
    gpu_ms = 0.0
    if mode == 'fp32+fp64':
        b32 = Batch(tenet, points, 'fp32s', estencil, rho0=rho0)
        if launch32:
            b32.set_launch(*launch32)
        for i, p in enumerate(points):
            b32.set_density(i, cold_start(p) if start is None else start[i])

        def restart32(i):
            b32.set_density(i, cold_start(points[i]))

        active = np.ones(n, dtype=bool)
        _, ms, fb = _relax(b32, active, steps, tol32, check_every, max_steps, grow_checks,
                           hist, restart32, 'fp32', plateau=True)
        gpu_ms += ms
        fb_all |= fb
        b64.copy_state_from(b32)
        b32.close()
    else:
        for i, p in enumerate(points):
            b64.set_density(i, cold_start(p) if start is None else start[i])

    def restart64(i):
        b64.set_density(i, cold_start(points[i]))

    active = np.ones(n, dtype=bool)
    rem, ms, fb = _relax(b64, active, steps, tol, check_every, max_steps, grow_checks, hist,
                         restart64, 'fp64')
    gpu_ms += ms
    fb_all |= fb
    fields = [b64.get_density(i) for i in range(n)]
    b64.close()
    wall = time.perf_counter() - t0
    return SweepResult(steps=steps, converged=~rem, fallback=fb_all, wall_s=wall,
                       gpu_ms=gpu_ms, history=hist), fields


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


def plain_runs(tenet, points, tol=1e-10, check_every=200, max_steps=400000, starts=None):
    """The unbatched baseline: one FP64 simulation per point, run the way a
    notebook runs MainLoop with a convergence callback, i.e. checkpoints of
    `check_every` steps and a device max |n - n_prev| check after each.
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
        fields.append(nprev.D2H().reshape(p.L, p.L))
        for x in (n, nprev, psi, u, a, b):
            x.Free()
        eng.close()
    return steps, conv, _t.perf_counter() - t0, fields
