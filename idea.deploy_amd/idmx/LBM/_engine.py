"""Thin handle around one idmx_sc_t (a Shan-Chen simulation in libidmx.so).

Only parameter marshalling and device pointers live here; every arithmetic
operation on lattice data runs in the HIP kernels.
"""
import contextlib
import ctypes as C

import numpy as np

from .. import _lib

XI_IDS = {'D2Q9': 0, 'D3Q19': 1}
E_IDS = {'E4': 4, 'E6': 6, 'E8': 8}
PSI_KINDS = {'exp': 0, 'expinv': 1, 'linear': 2, 'program': 3}
PRECISIONS = {'fp64': 0, 'fp32s': 1}

RUN_WRITE_FIELDS = 1
RUN_TWO_PASS = 2
RUN_FUSED_WALLS = 4


_KEEP = object()  # "leave this setting as it is" for SCEngine.configured


class SCEngine:
    def __init__(self, tenet, xi_name, e_name, dims, psi_kind, G, tau,
                 precision='fp64', rho0=0.0, mrt=None, forcing='guo', iso_filter=False,
                 eq_init=None, psi_code=None):
        self.tenet = tenet
        self.xi_name, self.e_name = xi_name, e_name
        self.dims = tuple(int(d) for d in dims)
        self.D = len(self.dims)
        self.Q = 9 if xi_name == 'D2Q9' else 19
        self.V = int(np.prod(self.dims))
        self.precision = precision
        self.psi_kind = psi_kind
        self.rho0 = float(rho0) if precision == 'fp32s' else 0.0
        self.pop_dtype = np.float64 if precision == 'fp64' else np.float32
        self.is_mrt = mrt is not None
        # NB x Q equilibrium coefficients for idmx_sc_init_pop_coeffs (None:
        # the compiled weights)
        self.eq_init = None if eq_init is None else \
            np.ascontiguousarray(eq_init, dtype=np.float64).ravel()
        if self.eq_init is not None and self.eq_init.size != (6 if self.D == 2 else 10) * self.Q:
            raise ValueError('eq_init must hold NB x Q = %d values'
                             % ((6 if self.D == 2 else 10) * self.Q))
        d = _lib.ScDesc()
        d.xi = XI_IDS[xi_name]
        d.estencil = E_IDS[e_name[2:]]
        for k in range(3):
            d.dims[k] = self.dims[k] if k < self.D else 1
        d.psi_kind = PSI_KINDS[psi_kind]
        d.SC_G = float(G)
        d.tau = float(tau)
        if forcing not in ('guo', 'sc') or (forcing == 'sc' and mrt is not None):
            raise ValueError("forcing is 'guo', or 'sc' (Shan-Chen forcing, BGK only)")
        if iso_filter and (forcing != 'guo' or mrt is not None or xi_name != 'D2Q9'
                           or precision != 'fp64'):
            raise ValueError('the isotropy filter is defined for D2Q9 BGK with the Guo '
                             'forcing in FP64')
        self.forcing = forcing
        self.iso_filter = bool(iso_filter)
        d.collision = (0 if forcing == 'guo' else 2) if mrt is None else 1
        if iso_filter:
            d.collision = 3
        d.precision = PRECISIONS[precision]
        d.rho0 = self.rho0
        keep = []
        if mrt is not None:
            if isinstance(mrt, dict) and 'diag' in mrt:
                a = np.ascontiguousarray(mrt['diag'], dtype=np.float64)
                keep.append(a)
                d.mrt_diag = a.ctypes.data
            else:
                for name, arr in zip(('mrt_A', 'mrt_beta', 'mrt_gamma'), mrt):
                    a = np.ascontiguousarray(arr, dtype=np.float64)
                    keep.append(a)
                    setattr(d, name, a.ctypes.data)
        h = C.c_void_p()
        _lib.call('idmx_sc_create', tenet.handle, C.byref(d), C.byref(h))
        self.handle = h
        tenet._live[id(self)] = self
        nb = C.c_size_t()
        _lib.call('idmx_sc_pop_bytes', h, C.byref(nb))
        self.pop_bytes = nb.value
        if psi_kind == 'program':
            if psi_code is None:
                raise ValueError("psi_kind 'program' needs psi_code")
            from .psi_code import compile_psi
            ops, consts = compile_psi(psi_code)
            _lib.call('idmx_sc_set_psi_program', h, ops.tobytes(), len(ops),
                      consts.ctypes.data_as(_lib.c_double_p), len(consts))
            self.psi_program = (ops, consts)

    def close(self):
        if self.handle:
            if self.tenet.handle:
                _lib.call('idmx_sc_destroy', self.handle)
            self.tenet._live.pop(id(self), None)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- kernels ----------------------------------------------------------
    def init_pop(self, pop, n, u):
        if self.eq_init is not None:
            _lib.call('idmx_sc_init_pop_coeffs', self.handle, pop.ptr, n.ptr, u.ptr,
                      self.eq_init.ctypes.data_as(_lib.c_double_p))
        else:
            _lib.call('idmx_sc_init_pop', self.handle, pop.ptr, n.ptr, u.ptr)

    def density_psi(self, n, psi, pop):
        _lib.call('idmx_sc_density_psi', self.handle, n.ptr, psi.ptr, pop.ptr)

    def force_collide_stream(self, pop_swap, pop, n, psi):
        _lib.call('idmx_sc_force_collide_stream', self.handle, pop_swap.ptr, pop.ptr,
                  n.ptr, psi.ptr)

    def run(self, pop, pop_swap, nsteps, n=None, psi=None, u=None,
            write_fields=False, two_pass=False, step_ms=None, fused_walls=False):
        """Advances nsteps; returns True when the state ended in pop_swap's
        allocation (the caller swaps its references, as M_SwapPop does).
        step_ms: a float32 array of 2 * nsteps values to receive each step's
        kernel times (idmx_sc_run_profiled)."""
        a = C.c_void_p(pop.ptr.value)
        b = C.c_void_p(pop_swap.ptr.value)
        flags = (RUN_WRITE_FIELDS if write_fields else 0) | (RUN_TWO_PASS if two_pass else 0) \
            | (RUN_FUSED_WALLS if fused_walls else 0)
        ptr = lambda x: x.ptr if x is not None else None
        if step_ms is None:
            _lib.call('idmx_sc_run', self.handle, C.byref(a), C.byref(b), ptr(n), ptr(psi),
                      ptr(u), int(nsteps), flags)
        else:
            if step_ms.dtype != np.float32 or step_ms.size < 2 * int(nsteps) \
                    or not step_ms.flags['C_CONTIGUOUS']:
                raise ValueError('step_ms must be a contiguous float32 array of 2 * nsteps')
            _lib.call('idmx_sc_run_profiled', self.handle, C.byref(a), C.byref(b), ptr(n),
                      ptr(psi), ptr(u), int(nsteps), flags,
                      step_ms.ctypes.data_as(C.POINTER(C.c_float)))
        return a.value != pop.ptr.value

    def step_kernel_names(self, two_pass=False):
        """Names of the (first, second) kernel of one step for per-step
        profiles: the two-pass step's density/psi and force-collide-stream
        kernels, or the single fused kernel."""
        body = any(v != 0.0 for v in getattr(self, '_body', (0.0, 0.0, 0.0)))
        walls = getattr(self, '_walls', None) is not None
        if (walls and not two_pass and self.precision == 'fp64' and not self.is_mrt
                and self.e_name.endswith('E4')
                and self.forcing == 'guo' and not self.iso_filter
                and self.psi_kind != 'program' and self.V >= (1 << 19)):
            # the fused step with walls: one kernel per step, the density /
            # psi kernel only in the run's last (two-pass) step
            return ('K_ComputeDensityPsiMeta', 'K_SCStep')
        if two_pass or body or walls or self.psi_kind == 'program':
            return ('K_ComputeDensityPsiMeta',
                    'K_ForceCollideStreamSCMP_MRTMeta' if self.is_mrt
                    else 'K_ForceCollideStreamSCMPMeta')
        return (None, 'K_SCStep')

    def moments(self, pop, n, psi, u):
        _lib.call('idmx_sc_moments', self.handle, pop.ptr, n.ptr, psi.ptr, u.ptr)

    def density_velocity(self, pop, n, u):
        _lib.call('idmx_sc_density_velocity', self.handle, pop.ptr, n.ptr, u.ptr)

    def mass_momentum(self, pop):
        m = C.c_double()
        mom = (C.c_double * 3)()
        _lib.call('idmx_sc_mass_momentum', self.handle, pop.ptr, C.byref(m), mom)
        return m.value, np.array(mom[:self.D])

    def last_run_ms(self):
        ms = C.c_double()
        _lib.call('idmx_sc_last_run_ms', self.handle, C.byref(ms))
        return ms.value

    def set_launch(self, tx, ty, H, hist=1):
        _lib.call('idmx_sc_set_launch', self.handle, int(tx), int(ty), int(H), int(hist))

    def set_walls(self, walls):
        """walls: device array of V uint8 flags (1 fluid, 0 solid), or None.
        While set, run / moments step as MainLoopWalls."""
        _lib.call('idmx_sc_set_walls', self.handle, walls.ptr if walls is not None else None)
        self._walls = walls

    def set_body_force(self, fx=0.0, fy=0.0, fz=0.0):
        """Body force added to the Shan-Chen force (gravity of
        MainLoopSimpleWallsGravity); zero restores the plain step."""
        _lib.call('idmx_sc_set_body_force', self.handle, float(fx), float(fy), float(fz))
        self._body = (float(fx), float(fy), float(fz))

    @contextlib.contextmanager
    def configured(self, walls=_KEEP, body=_KEEP):
        """Walls and / or body force for the duration of a with-block; the
        previous settings are restored afterwards (so a convergence callback
        that computes moments inside a wall loop leaves the loop's walls in
        place)."""
        old_w = getattr(self, '_walls', None)
        old_b = getattr(self, '_body', (0.0, 0.0, 0.0))
        try:
            if walls is not _KEEP:
                self.set_walls(walls)
            if body is not _KEEP:
                self.set_body_force(*body)
            yield self
        finally:
            self.set_walls(old_w)
            self.set_body_force(*old_b)

    def launch_ok(self, tx, ty, H, hist=1):
        """Whether set_launch accepts this shape (variant 9 needs its LDS
        staging buffer to fit one CU); the current launch is kept."""
        old = self.get_launch()
        try:
            self.set_launch(tx, ty, H, hist)
            return True
        except _lib.IdmxError:
            return False
        finally:
            self.set_launch(*old)

    def get_launch(self):
        v = [C.c_int() for _ in range(4)]
        _lib.call('idmx_sc_get_launch', self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    # -- host views of a population buffer ---------------------------------
    def pop_weights(self):
        from .stencils import XIStencils
        return np.array(XIStencils[self.xi_name]['Ws'], dtype=np.float64)

    def pop_to_host_soa(self, pop):
        """float64 SoA populations (Q, V) with the shift removed."""
        from ..IdpyCode.IdpyMemory import IdpyArrayHIP
        raw = IdpyArrayHIP.D2H(pop).reshape(self.Q, self.V).astype(np.float64)
        if self.precision == 'fp32s':
            raw += self.pop_weights()[:, None] * self.rho0
        return raw

    def pop_from_host_soa(self, pop, f):
        f = np.asarray(f, dtype=np.float64).reshape(self.Q, self.V)
        if self.precision == 'fp32s':
            f = f - self.pop_weights()[:, None] * self.rho0
        from ..IdpyCode.IdpyMemory import IdpyArrayHIP
        IdpyArrayHIP.H2D(pop, f.astype(self.pop_dtype))
