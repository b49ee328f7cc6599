"""ctypes wrapper of the serial C++ oracle libcpuref.so (TEST INFRASTRUCTURE).

Builds nothing: oracle/Makefile makes libcpuref.so (build() of
__graft_entry__ runs it). Arrays are numpy float64, C-contiguous.
"""
import ctypes as C
import os

import numpy as np

from .stencils import E, XI

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libcpuref.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/libcpuref.so is not built (run make -C oracle)')
        L = C.CDLL(path)
        L.cpuref_legacy_run.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                        C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                        C.c_int]
        L.cpuref_meta_run.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, C.c_int,
                                      C.c_void_p]
        L.cpuref_meta_run_mom.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                          C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                          _dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.cpuref_meta_moments.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                          C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_void_p]
        L.cpuref_legacy_walls_run.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                              C.c_double, C.c_int, C.c_void_p, C.c_double,
                                              C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.cpuref_iso_run.argtypes = [_ip, C.c_int, _ip, _dp, C.c_double, C.c_double,
                                     C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.c_int]
        L.cpuref_iso_filter.argtypes = [C.c_int64, _dp, _dp, _dp]
        L.cpuref_init_pop.argtypes = [C.c_int, _ip, C.c_int, _dp, _dp, _dp]
        L.cpuref_set_psi_fn.argtypes = [C.c_void_p]
        L.cpuref_meta_run_terms.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                            C.c_int, C.c_int, _ip, _dp, _ip, _dp, _dp, _dp,
                                            _dp, C.c_int]
        L.cpuref_init_pop_terms.argtypes = [C.c_int, _ip, C.c_int, _dp, _dp, _ip, _dp, _ip,
                                            _dp]
        L.cpuref_meta_run_dense.argtypes = [C.c_int, _ip, C.c_int, _ip, _dp, C.c_double,
                                            C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                            _dp, C.c_int]
        L.cpuref_init_radial.argtypes = [C.c_int, _ip, C.c_double, C.c_double,
                                         C.c_double, C.c_double, _dp, _dp]
        L.cpuref_init_flat.argtypes = [C.c_int, _ip, C.c_double, C.c_double, C.c_double,
                                       C.c_int, C.c_int, _dp, _dp]
        L.cpuref_init_cylinder.argtypes = [C.c_int, _ip, C.c_double, C.c_double, C.c_double,
                                           C.c_int, C.c_int, _dp, _dp]
        L.cpuref_init_single_flat.argtypes = [C.c_int, _ip, C.c_double, C.c_double, C.c_int,
                                              C.c_int, _dp, _dp]
        _LIB = L
    return _LIB


def compile_psi_code(psi_code, workdir, var='ln_0'):
    """Compiles the C expression psi_code into double psi(double var) with
    gcc (-O3, no floating-point contraction, as the CTypes backend builds
    K_ComputeDensityPsiMeta) and makes it the oracle's psi kind 'code'.
    Returns the loaded library (keep it alive while the oracle runs)."""
    import subprocess
    src = os.path.join(workdir, 'psi_code.c')
    so = os.path.join(workdir, 'psi_code.so')
    with open(src, 'w') as fh:
        fh.write('#include <math.h>\n'
                 'typedef double NType; typedef double PsiType;\n'
                 'double psi(double %s) { return %s; }\n' % (var, psi_code))
    subprocess.run(['gcc', '-std=c99', '-O3', '-ffp-contract=off', '-fPIC', '-shared', '-o', so,
                    src, '-lm'], check=True)
    L = C.CDLL(so)
    fn = C.cast(L.psi, C.c_void_p).value
    lib().cpuref_set_psi_fn(C.c_void_p(fn))
    return L


def _xi_id(xi):
    return {'D2Q9': 0, 'D3Q19': 1}[xi]


def _estencil(estencil):
    es, ws = E[estencil]
    Ei = np.ascontiguousarray(np.array(es, dtype=np.int32).ravel())
    Ew = np.ascontiguousarray(np.array([float(w) for w in ws], dtype=np.float64))
    return len(es), Ei, Ew


def _L(dims):
    return np.ascontiguousarray(np.array(list(dims) + [1] * (3 - len(dims)), dtype=np.int32))


PSI_KIND = {'exp': 0, 'expinv': 1, 'linear': 2, 'code': 3}


class State:
    """Fields of one oracle simulation (pop in the given ordering)."""

    def __init__(self, xi, dims, aos=False):
        self.xi, self.dims, self.aos = xi, tuple(dims), aos
        self.Q = len(XI[xi])
        self.D = len(dims)
        self.V = int(np.prod(dims))
        self.pop = np.zeros(self.V * self.Q)
        self.pop_swap = np.zeros(self.V * self.Q)
        self.n = np.zeros(self.V)
        self.psi = np.zeros(self.V)
        self.u = np.zeros(self.V * self.D)


def init_radial(dims, n_g, n_l, R, W):
    """Legacy K_InitRadialInterface (idpy/LBM/LBM.py:1786-1826)."""
    V = int(np.prod(dims))
    n = np.zeros(V)
    u = np.zeros(V * len(dims))
    rc = lib().cpuref_init_radial(len(dims), _L(dims), n_g, n_l, R, W, n, u)
    assert rc == 0
    return n, u


def init_flat(dims, n_g, n_l, width, direction=0, full_flag=True):
    """K_InitFlatInterface (idpy/LBM/LBM.py:1745-1784)."""
    V = int(np.prod(dims))
    n, u = np.zeros(V), np.zeros(V * len(dims))
    assert lib().cpuref_init_flat(len(dims), _L(dims), n_g, n_l, width, direction,
                                  int(bool(full_flag)), n, u) == 0
    return n, u


def init_cylinder(dims, n_g, n_l, R, direction=0, full_flag=True):
    """K_InitCylinderInterface (idpy/LBM/LBM.py:1828-1874)."""
    V = int(np.prod(dims))
    n, u = np.zeros(V), np.zeros(V * len(dims))
    assert lib().cpuref_init_cylinder(len(dims), _L(dims), n_g, n_l, R, direction,
                                      int(bool(full_flag)), n, u) == 0
    return n, u


def init_single_flat(dims, n_g, n_l, direction=0, full_flag=True):
    """K_InitSingleFlatInterface (idpy/LBM/MultiPhaseKernels.py:650-684)."""
    V = int(np.prod(dims))
    n, u = np.zeros(V), np.zeros(V * len(dims))
    assert lib().cpuref_init_single_flat(len(dims), _L(dims), n_g, n_l, direction,
                                         int(bool(full_flag)), n, u) == 0
    return n, u


def init_pop(st, n, u):
    """K_InitPopulations (idpy/LBM/LBM.py:1700-1743); u is SoA."""
    rc = lib().cpuref_init_pop(_xi_id(st.xi), _L(st.dims), int(st.aos),
                               np.ascontiguousarray(n, dtype=np.float64),
                               np.ascontiguousarray(u, dtype=np.float64), st.pop)
    assert rc == 0


def init_pop_ref15(st, n, u):
    """K_InitPopulationsMeta with the current class's printed 15-digit
    coefficients (oracle/ref15_terms.py); u is SoA."""
    from . import ref15_terms
    off, coef, fac = ref15_terms.equilibrium_terms(st.xi)
    rc = lib().cpuref_init_pop_terms(_xi_id(st.xi), _L(st.dims), int(st.aos),
                                     np.ascontiguousarray(n, dtype=np.float64),
                                     np.ascontiguousarray(u, dtype=np.float64), off, coef,
                                     fac, st.pop)
    assert rc == 0


def meta_run_ref15(st, estencil, G, tau, psi='exp', nsteps=1):
    """Meta class MainLoop (idpy/LBM/MultiPhase.py:1015-1097) with the
    collision evaluated as the printed sums of K_ForceCollideStreamSCMPMeta
    (15-digit coefficients, oracle/ref15_terms.py)."""
    from . import ref15_terms
    QE, Ei, Ew = _estencil(estencil)
    off, coef, fac = ref15_terms.collision_terms(st.xi, 1.0 / tau)
    rc = lib().cpuref_meta_run_terms(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, PSI_KIND[psi],
                                     int(st.aos), off, coef, fac, st.pop, st.pop_swap, st.n,
                                     st.psi, nsteps)
    assert rc == 0


def meta_run_dense_tables(st, estencil, G, tables, psi='exp', nsteps=1):
    """Meta class MainLoop(MRT) with the collision as the dense
    basis-coordinate tables (A Q x Q, beta and gamma NB x Q) that the
    library's dense kernel takes."""
    QE, Ei, Ew = _estencil(estencil)
    A, beta, gamma = (np.ascontiguousarray(t, dtype=np.float64) for t in tables)
    rc = lib().cpuref_meta_run_dense(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, PSI_KIND[psi],
                                     int(st.aos), A, beta, gamma, st.pop, st.pop_swap, st.n,
                                     st.psi, nsteps)
    assert rc == 0


def legacy_run(st, estencil, G, tau, psi='exp', nsteps=1):
    """Legacy class MainLoop, nsteps (idpy/LBM/LBM.py:806-894). SoA only."""
    assert not st.aos
    QE, Ei, Ew = _estencil(estencil)
    rc = lib().cpuref_legacy_run(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, 1.0 / tau,
                                 PSI_KIND[psi], st.pop, st.pop_swap, st.n, st.u,
                                 st.psi, nsteps)
    assert rc == 0


def legacy_walls_run(st, estencil, G, tau, walls, g=0.0, psi='exp', nsteps=1, forcing='guo'):
    """MainLoopSimpleWalls (g = 0) / MainLoopSimpleWallsGravity
    (idpy/LBM/MultiPhase.py:351-465, 888-1010), or with forcing='sc'
    MainLoopSimpleWallsSC (:467-581; walls None: MainLoopSimpleSC,
    :257-349), nsteps; SoA; st.psi must hold the solid sites' psi
    (GetWalls)."""
    assert not st.aos and forcing in ('guo', 'sc')
    QE, Ei, Ew = _estencil(estencil)
    wkeep, wptr = _walls_ptr(walls, st.V)
    rc = lib().cpuref_legacy_walls_run(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, 1.0 / tau,
                                       PSI_KIND[psi], wptr, float(g), int(forcing == 'sc'),
                                       st.pop, st.pop_swap, st.n, st.u, st.psi, nsteps)
    assert rc == 0


def _walls_ptr(walls, V):
    if walls is None:
        return None, None
    w = np.ascontiguousarray(walls, dtype=np.uint8).ravel()
    assert w.size == V
    return w, w.ctypes.data


def meta_run(st, estencil, G, tau, psi='exp', nsteps=1, mrt=None, walls=None):
    """Meta class MainLoop / MainLoopMRT, nsteps
    (idpy/LBM/MultiPhase.py:1015-1097, 1311-1402). mrt = (A, B, Gg). walls
    (V flags, 1 fluid / 0 solid, lexicographic): the step of MainLoopWalls
    (MultiPhase.py:1226-1309); st.psi must then hold the solid sites' psi."""
    QE, Ei, Ew = _estencil(estencil)
    wkeep, wptr = _walls_ptr(walls, st.V)
    ptrs = [None, None, None]
    keep = []
    if mrt is not None:
        for i, m in enumerate(mrt):
            a = np.ascontiguousarray(m, dtype=np.float64)
            keep.append(a)
            ptrs[i] = a.ctypes.data
    rc = lib().cpuref_meta_run(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, 1.0 / tau,
                               PSI_KIND[psi], int(st.aos), ptrs[0], ptrs[1], ptrs[2],
                               st.pop, st.pop_swap, st.n, st.psi, nsteps, wptr)
    assert rc == 0


def meta_run_mrt_moment(st, estencil, G, rates, psi='exp', nsteps=1):
    """MainLoopMRT with the diagonal relaxation matrix diag(rates)
    (idpy/LBM/MultiPhase.py:1311-1402), the collision evaluated in moment
    space through the lattice monomials (same result as meta_run with the
    dense matrices, up to rounding; the same sums in the same order as the
    HIP kernel's moment-space collision)."""
    from . import hermite_mrt
    QE, Ei, Ew = _estencil(estencil)
    t = hermite_mrt.moment_tables(st.xi)
    s = np.ascontiguousarray(np.asarray(rates, dtype=np.float64).ravel())
    assert s.size == (9 if st.xi == 'D2Q9' else 19)
    rc = lib().cpuref_meta_run_mom(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G, PSI_KIND[psi],
                                   int(st.aos), len(t['pairs']), t['pairs'], t['mm'], t['odd'],
                                   t['tm'], t['lt'], t['dinv'], t['ht'], t['mpv'], s, st.pop,
                                   st.pop_swap, st.n, st.psi, nsteps)
    assert rc == 0


def meta_moments(st, estencil, G, psi='exp', walls=None):
    """ComputeMomentsMeta (idpy/LBM/MultiPhase.py:1565-1662): fills n, psi, u
    with u = sum xi f / n + F / (2n); u in the same ordering as pop. With
    walls only the fluid sites are written."""
    QE, Ei, Ew = _estencil(estencil)
    wkeep, wptr = _walls_ptr(walls, st.V)
    rc = lib().cpuref_meta_moments(_xi_id(st.xi), _L(st.dims), QE, Ei, Ew, G,
                                   PSI_KIND[psi], int(st.aos), st.pop, st.n, st.u, st.psi,
                                   wptr)
    assert rc == 0


def _iso_tables():
    from . import iso_filter
    M, Ai = iso_filter.tables()
    return np.ascontiguousarray(M), np.ascontiguousarray(Ai)


def iso_run(st, estencil, G, tau, psi='exp', nsteps=1):
    """MainLoopIsoFilter (idpy/LBM/MultiPhase.py:1099-1223), D2Q9 BGK:
    the Meta step with K_IsotropyFilter between collision and push."""
    assert st.xi == 'D2Q9'
    QE, Ei, Ew = _estencil(estencil)
    M, Ai = _iso_tables()
    rc = lib().cpuref_iso_run(_L(st.dims), QE, Ei, Ew, G, 1.0 / tau, PSI_KIND[psi],
                              int(st.aos), M, Ai, st.pop, st.pop_swap, st.n, st.psi, nsteps)
    assert rc == 0


def iso_filter(f):
    """K_IsotropyFilter (idpy/LBM/LBMKernelsMeta.py:895-1080) on an
    (nsites, 9) array of populations; returns the filtered copy."""
    out = np.ascontiguousarray(f, dtype=np.float64).copy()
    assert out.ndim == 2 and out.shape[1] == 9
    M, Ai = _iso_tables()
    rc = lib().cpuref_iso_filter(out.shape[0], M, Ai, out)
    assert rc == 0
    return out
