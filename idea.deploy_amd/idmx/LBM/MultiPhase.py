"""idpy.LBM.MultiPhase surface: the current ShanChenMultiPhase class
(idpy/LBM/MultiPhase.py:140-2175) on the HIP backend.

MainLoop / MainLoopMRT run the fused Shan-Chen step (density, psi, E-stencil
force, Guo-shifted velocity, BGK or MRT collision with Guo forcing, push
streaming) on the device. Observable semantics follow the reference: after
MainLoop, n and psi hold the values of the start of the last step and u is
not written (call ComputeMomentsMeta first, MultiPhaseLoopChecks.py:125).
"""
import numpy as np

from ..IdpyCode import HIP_T  # noqa: F401
from ..IdpyCode import IdpyMemory
from ._base import (AllTrue, IndexFromPos, OneTrue, PosFromIndex,  # noqa: F401
                    _ShanChenCommon)
from .stencils import FStencils, XIStencils  # noqa: F401


def ComputeCenterOfMass(lbm, c_i=''):
    """(mass, centre of mass) of n (idpy/LBM/MultiPhase.py:108-137)."""
    return lbm.ComputeCenterOfMass()


class ShanChenMultiPhase(_ShanChenCommon):
    """Drop-in for idpy.LBM.MultiPhase.ShanChenMultiPhase.

    kwargs as in SetupRoot (MultiPhase.py:2096-2108): lang, device,
    custom_types, block_size, f_stencil, psi_code (C expression in ln_0),
    SC_G, tau, optimizer_flag, e2_val, psi_sym, empty_sim, set_ordering,
    use_ptrs, fp32_flag, dim_sizes, xi_stencil; plus precision ('fp64' |
    'fp32s'), rho0 and step_mode ('fused' | 'two_pass'). fp32_flag selects
    'fp32s'. The fluctuating-LBM kwargs are not supported.

    constants: 'exact' (default; full-double coefficients, mass conserved to
    round-off, the arithmetic of the legacy class) or 'ref15' (the 15-digit
    coefficients this class's generated kernels print, evalf() / str() of the
    sympy expansion, idpy/LBM/Collisions.py:251-273; with tau = 1 the D2Q9
    step then loses 8.5e-16 of the mass per step, as the reference does).
    'ref15' applies to InitPopulations, MainLoop, MainLoopMRT and
    MainLoopWalls, which then run the dense collision with tables holding
    the printed coefficients; FP64 only.
    """
    _PRINTED_CONSTANTS = True

    _NEEDED = ['lang', 'cl_kind', 'device', 'custom_types', 'block_size', 'f_stencil',
               'psi_code', 'SC_G', 'tau', 'optimizer_flag', 'e2_val', 'psi_sym',
               'empty_sim', 'set_ordering', 'use_ptrs', 'fluctuations', 'prng_kind',
               'init_seed', 'prng_init_from', 'prng_distribution', 'indep_gaussian',
               'fp32_flag']

    def __init__(self, *args, **kwargs):
        self._setup(self._NEEDED, args, kwargs, 'ln_0')
        if self.params_dict.get('fluctuations') is not None:
            raise NotImplementedError('fluctuating LBM (Gross 2011) is not supported')
        so = self.params_dict.get('set_ordering', None)
        if so not in ('gpu', 'cpu', None):
            raise Exception("Parameter 'set_ordering' must either be 'gpu' or 'cpu'")
        self.params_dict['set_ordering'] = so if so is not None else self.tenet.GetKind()
        self.params_dict.setdefault('use_ptrs', True)
        self._init_vars()
        self.sims_vars['ordering'] = \
            self.sims_vars['ordering_lambdas'][self.params_dict['set_ordering']]
        for name in ['n', 'u']:
            self.sims_idpy_memory[name] = None
        if not self.params_dict['empty_sim']:
            aos = self.params_dict['set_ordering'] == 'cpu'
            self._init_memory(pop_aos=aos, u_aos=aos)
            # the step's own velocity goes to a private buffer: MainLoop does
            # not write 'u' on the reference
            self._u_step = IdpyMemory.Zeros(self.sims_vars['V'] * self.sims_vars['DIM'],
                                            np.float64, self.tenet)
        self.init_status = {'n': False, 'u': False, 'pop': False}

    def MainLoop(self, time_steps, convergence_functions=[], profiling=False):
        return self._run_loop(time_steps, convergence_functions, profiling, self.engine,
                              self._u_step)

    def MainLoopMRT(self, time_steps, relaxation_matrix=None, omega_syms_vals=None,
                    convergence_functions=[], profiling=False):
        """MRT time loop (MultiPhase.py:1311-1402). relaxation_matrix is the
        Q x Q matrix S in the W-orthogonal Hermite moment order (a sympy
        Matrix whose symbols omega_syms_vals = [(symbol, value), ...] fixes,
        or a numeric array)."""
        if relaxation_matrix is None:
            raise Exception("Missing argument 'relaxation_matrix'")
        if omega_syms_vals is None:
            raise Exception("Missing argument 'omega_syms_vals'")
        eng = self._mrt_engine(relaxation_matrix, omega_syms_vals)
        return self._run_loop(time_steps, convergence_functions, profiling, eng,
                              self._u_step)

    def MainLoopWalls(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopWalls (MultiPhase.py:1226-1309): the Meta step on the fluid
        sites of the walls set by GetWalls, push streaming with half-way
        bounce-back at the solid sites, whose psi (GetWalls' psi_w) enters
        the force of their fluid neighbours. Runs the two-pass step."""
        m = self.sims_idpy_memory
        if 'walls' not in m:
            raise Exception("Missing 'walls': call GetWalls first")
        with self.engine.configured(walls=m['walls']):
            return self._run_loop(time_steps, convergence_functions, profiling, self.engine,
                                  self._u_step)

    def MainLoopIsoFilter(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopIsoFilter (MultiPhase.py:1099-1223), D2Q9: the Meta step
        with the isotropy filter (K_IsotropyFilter, LBMKernelsMeta.py:
        895-1080) applied to every site's post-collision populations before
        the push stream. FP64 only (the filter is written for nine FP64
        populations). Runs the two-pass step; u is not written."""
        if self.xi_name != 'D2Q9':
            raise NotImplementedError('MainLoopIsoFilter is defined for D2Q9')
        if self.params_dict['precision'] != 'fp64':
            raise NotImplementedError('MainLoopIsoFilter runs in FP64')
        return self._run_loop(time_steps, convergence_functions, profiling,
                              self._iso_engine(), self._u_step)

    def MainLoopSimpleSC(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopSimpleSC (MultiPhase.py:257-349): BGK with the Shan-Chen
        forcing (K_Collision_ShanChenMultiPhase, MultiPhaseKernels.py:132-216:
        equilibrium velocity u + F/(n omega), no Guo term), pull streaming.
        Writes n, psi and u = (u + u_post) / 2 as the legacy kernels do."""
        m = self.sims_idpy_memory
        return self._run_loop(time_steps, convergence_functions, profiling,
                              self._scf_engine(), m['u'])

    def MainLoopSimpleWallsSC(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopSimpleWallsSC (MultiPhase.py:467-581): MainLoopSimpleSC on
        the fluid sites of the walls set by GetWalls, with half-way
        bounce-back (K_Collision_ShanChenMultiPhaseWalls,
        MultiPhaseKernels.py:218-304; K_HalfWayBounceBack)."""
        m = self.sims_idpy_memory
        if 'walls' not in m:
            raise Exception("Missing 'walls': call GetWalls first")
        eng = self._scf_engine()
        with eng.configured(walls=m['walls']):
            return self._run_loop(time_steps, convergence_functions, profiling, eng, m['u'])

    def MainLoopSimple(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopSimple (MultiPhase.py:164-255): the legacy-kernel loop
        (K_ComputeMoments, K_ComputePsi, K_Collision_ShanChenGuoMultiPhase,
        K_StreamPeriodic, LBM.py:1529-1698) on the periodic lattice: the same
        update as MainLoop, with n, psi and the Guo-shifted u of the last step
        written, as the legacy kernels leave them."""
        return self._run_loop(time_steps, convergence_functions, profiling, self.engine,
                              self.sims_idpy_memory['u'])

    def MainLoopSimpleWallsInletOutlet(self, time_steps, convergence_functions=[],
                                       profiling=False, n_in=1, u_in=0, tau_in=100, max_mult=2):
        """Not provided. MultiPhase.py:583-733 deploys K_SetPopNUXInletDutyCycle
        with 7 arguments (:659) against the 9 parameters of its signature
        (LBMKernels.py:307-315), and the kernel sets n = 2 whatever n_in is
        (LBMKernels.py:330), so the reference defines no behaviour to match."""
        raise NotImplementedError('MainLoopSimpleWallsInletOutlet: the reference loop passes 7 '
                                  'arguments to a 9-parameter inlet kernel; not provided')

    def MainLoopSimpleWallsInletOutletSC(self, time_steps, convergence_functions=[],
                                         profiling=False, n_in=1, u_in=0, tau_in=100,
                                         max_mult=2):
        """Not provided, for the reason given in MainLoopSimpleWallsInletOutlet
        (the same inlet kernel call, MultiPhase.py:735-886)."""
        raise NotImplementedError('MainLoopSimpleWallsInletOutletSC: the reference loop passes 7 '
                                  'arguments to a 9-parameter inlet kernel; not provided')

    def MainLoopGross2011SRT(self, *args, **kwargs):
        """Not provided: fluctuating LBM (MultiPhase.py:1404-1541) is outside
        the deterministic step this library implements."""
        raise NotImplementedError('MainLoopGross2011SRT: fluctuating LBM is not supported')

    def ComputeMoments(self):
        """ComputeMoments (MultiPhase.py:1543-1563): K_ComputeMoments on the
        device, n = sum f and u = sum xi f / n (no force shift)."""
        if not self.init_status['pop']:
            raise Exception('Populations are not initialized')
        m = self.sims_idpy_memory
        self.engine.density_velocity(m['pop'], m['n'], m['u'])
        self.init_status['n'] = True
        self.init_status['u'] = True

    def MainLoopSimpleWalls(self, time_steps, convergence_functions=[], profiling=False):
        """MainLoopSimpleWalls (MultiPhase.py:351-465): the legacy-kernel walls
        loop (K_ComputeMomentsWalls, K_ComputePsiWalls,
        K_Collision_ShanChenGuoMultiPhaseWalls, K_StreamPeriodic,
        K_HalfWayBounceBack). Fluid sites follow the same update as
        MainLoopWalls; u (Guo-shifted), n and psi are written as by the
        legacy kernels. The solid sites' populations are not defined (the
        reference leaves them order-dependent)."""
        return self.MainLoopSimpleWallsGravity(time_steps, convergence_functions, g=0,
                                               profiling=profiling)

    def MainLoopSimpleWallsGravity(self, time_steps, convergence_functions=[], g=0,
                                   profiling=False):
        """MainLoopSimpleWallsGravity (MultiPhase.py:888-1010): as
        MainLoopSimpleWalls with the gravity g added to the y component of the
        force (F[1] += g, MultiPhaseKernels.py:474-560)."""
        m = self.sims_idpy_memory
        if 'walls' not in m:
            raise Exception("Missing 'walls': call GetWalls first")
        with self.engine.configured(walls=m['walls'], body=(0.0, float(g), 0.0)):
            return self._run_loop(time_steps, convergence_functions, profiling, self.engine,
                                  m['u'])

    def GetWalls(self, walls, psi_w):
        """psi_w (V values, the psi of the solid sites; the fluid sites' are
        recomputed every step) and the walls flags (1 fluid, 0 solid, shape
        flip(dim_sizes)) to the device (MultiPhase.py:157-161, LBM.py:609-631)."""
        V = self.sims_vars['V']
        psi_w = np.ravel(np.asarray(psi_w, dtype=np.float64))
        if psi_w.size != V:
            raise ValueError('psi_w must hold V = %d values' % V)
        self.sims_idpy_memory['psi'].H2D(psi_w)
        m = self.sims_idpy_memory
        if 'xi_opposite' not in m:
            xi = np.array(self.params_dict['xi_stencil']['XIs'])
            opp = [int(np.where((xi == -c).all(axis=1))[0][0]) for c in xi]
            m['xi_opposite'] = IdpyMemory.OnDevice(np.array(opp, dtype=np.int32), self.tenet)
        if 'walls' not in m:
            m['walls'] = IdpyMemory.OnDevice(np.ones(V, dtype=np.uint8), self.tenet)
        dims_rev = list(self.sims_vars['dim_sizes'])[::-1]
        walls = np.asarray(walls)
        if list(walls.shape) == dims_rev:
            m['walls'].H2D(np.ravel(walls).astype(np.uint8))

    def ComputeMomentsMeta(self):
        """n, psi and the Guo-shifted u = sum xi f / n + F/(2n) on the device
        (MultiPhase.py:1565-1662); with walls, on the fluid sites only."""
        if not self.init_status['pop']:
            raise Exception('Populations are not initialized')
        m = self.sims_idpy_memory
        # as on the reference: the walls whenever they exist, never a body
        # force (the Meta moments kernels have none); a wall loop calling this
        # from a convergence callback gets its own settings back afterwards
        with self.engine.configured(walls=m.get('walls'), body=(0.0, 0.0, 0.0)):
            self.engine.moments(m['pop'], m['n'], m['psi'], m['u'])
        self.init_status['n'] = True
        self.init_status['u'] = True

    def InitRadialInterface(self, n_g, n_l, R, full_flag=True):
        """K_InitRadialInterface, tanh(r - R) (MultiPhaseKernels.py:686-727),
        on the device."""
        self._init_radial(n_g, n_l, R, 1.0, full_flag)

    def InitSingleFlatInterface(self, n_g, n_l, direction=0, full_flag=True):
        """K_InitSingleFlatInterface with F_NSingleFlatProfile
        (MultiPhaseKernels.py:45-53, 650-684), on the device."""
        self.sims_vars.update({'init_type': 'flat', 'n_g': n_g, 'n_l': n_l,
                               'full_flag': full_flag, 'direction': direction})
        self._device_init('idmx_init_single_flat', float(n_g), float(n_l), int(direction),
                          int(bool(full_flag)))
        self.InitPopulations()
