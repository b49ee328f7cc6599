"""Legacy idpy.LBM.LBM surface: stencil tables, LBMTypes, RootLB and the
ShanChenMultiPhase class of idpy/LBM/LBM.py:788-1279.

The legacy class runs the step as four kernels (moments, psi, collision,
pull streaming) with full-double weights read from memory; it conserves mass
to round-off. Here the same step runs as one fused HIP kernel per time step.
Its observable state after MainLoop is kept: n and psi hold the values
computed at the start of the last step, u holds that step's Guo-shifted
velocity (K_Collision_ShanChenGuoMultiPhase writes it, LBM.py:1577-1580).
"""
from ..IdpyCode import HIP_T  # noqa: F401
from ._base import (AllTrue, CustomTypes, IndexFromPos, LBMTypes, OneTrue,  # noqa: F401
                    PosFromIndex, RootLB, _ShanChenCommon)
from .stencils import FStencils, XIStencils  # noqa: F401


class ShanChenMultiPhase(_ShanChenCommon):
    """Drop-in for idpy.LBM.LBM.ShanChenMultiPhase (LBM.py:788-1279).

    kwargs as in SetupRoot (LBM.py:1227-1232): lang, device, custom_types,
    block_size, f_stencil, psi_code (C expression in ln), SC_G, tau,
    optimizer_flag, e2_val, psi_sym, empty_sim, dim_sizes, xi_stencil; plus
    precision ('fp64' | 'fp32s'), rho0 and step_mode ('fused' | 'two_pass').
    """

    _NEEDED = ['lang', 'cl_kind', 'device', 'custom_types', 'block_size', 'f_stencil',
               'psi_code', 'SC_G', 'tau', 'optimizer_flag', 'e2_val', 'psi_sym',
               'empty_sim']

    def __init__(self, *args, **kwargs):
        self._setup(self._NEEDED, args, kwargs, 'ln')
        self._init_vars()
        for name in ['n', 'u']:
            self.sims_idpy_memory[name] = None
        if not self.params_dict['empty_sim']:
            self._init_memory(pop_aos=False, u_aos=False)
        self.init_status = {'n': False, 'u': False, 'pop': False}

    def MainLoop(self, time_steps, convergence_functions=[], profiling=False):
        return self._run_loop(time_steps, convergence_functions, profiling, self.engine,
                              self.sims_idpy_memory['u'])

    def ComputeMoments(self):
        """K_ComputeMoments (LBM.py:1667-1698) on the device: n = sum f,
        u = sum xi f / n (no force shift)."""
        if not self.init_status['pop']:
            raise Exception('Populations are not initialized')
        m = self.sims_idpy_memory
        self.engine.density_velocity(m['pop'], m['n'], m['u'])
        self.init_status['n'] = True
        self.init_status['u'] = True

    def InitRadialInterface(self, n_g, n_l, R, W=1, full_flag=True):
        """K_InitRadialInterface, tanh((r - R)/W) (LBM.py:1786-1826), on the
        device."""
        self._init_radial(n_g, n_l, R, W, full_flag)
