"""Convergence callbacks run between MainLoop checkpoints
(idpy/LBM/MultiPhaseLoopChecks.py:37-163), on device reductions."""
import ctypes as C

import numpy as np

from .. import _lib
from ..IdpyCode import IdpyMemory


def CheckUConvergenceSCMP(lbm, skip_exit_condition=False, print_flag=True):
    """mean |delta u| and max Mach sqrt(max u^2 / c2) between checkpoints
    (MultiPhaseLoopChecks.py:96-163); True once mean |delta u| < 1e-12
    (1e-5 for single-precision velocities)."""
    first_flag = False
    m = lbm.sims_idpy_memory
    if 'old_u' not in m:
        m['old_u'] = IdpyMemory.Zeros(lbm.sims_vars['V'] * lbm.sims_vars['DIM'],
                                      np.float64, lbm.tenet)
        lbm.aux_idpy_memory.append('old_u')
        lbm.sims_vars['u_conv'], lbm.sims_vars['max_u'] = [], []
        lbm.aux_vars += ['u_conv', 'max_u']
        first_flag = True
    lbm.ComputeMomentsMeta()
    sd, mu = C.c_double(), C.c_double()
    _lib.call('idmx_check_u', lbm.tenet.handle, m['u'].ptr, m['old_u'].ptr,
              lbm.sims_vars['V'], lbm.sims_vars['DIM'], C.byref(sd), C.byref(mu))
    u_conv = sd.value / lbm.sims_vars['V']
    max_u = np.sqrt(mu.value / lbm.sims_vars['c2'])
    lbm.sims_vars['u_conv'].append(u_conv)
    lbm.sims_vars['max_u'].append(max_u)
    _u_threshold = 1e-12
    if print_flag:
        print('u_conv: ', u_conv, 'max_u: ', max_u)
        print('Conv!', u_conv < _u_threshold, first_flag)
    if skip_exit_condition:
        return False
    return (not first_flag) and u_conv < _u_threshold


def CheckCenterOfMassDeltaPConvergence(lbm):
    """Laplace pressure and centre of mass history; True when the droplet
    drifts off centre or Delta p stops changing (MultiPhaseLoopChecks.py:37-89)."""
    first_flag = False
    if 'delta_p' not in lbm.sims_vars:
        lbm.sims_vars['delta_p'], lbm.sims_vars['p_in'], lbm.sims_vars['p_out'] = [], [], []
        lbm.sims_vars['cm_conv'] = []
        lbm.aux_vars += ['delta_p', 'p_in', 'p_out', 'cm_conv']
        first_flag = True
    lbm.ComputeMomentsMeta()
    p_in, p_out, dp = lbm.DeltaPLaplace()
    lbm.sims_vars['delta_p'].append(dp)
    lbm.sims_vars['p_in'].append(p_in)
    lbm.sims_vars['p_out'].append(p_out)
    lbm.sims_vars['cm_conv'].append(lbm.sims_vars['cm_coords'])
    if not lbm.sims_vars['is_centered']:
        return True
    if first_flag or len(lbm.sims_vars['delta_p']) < 2:
        return False
    d = abs(lbm.sims_vars['delta_p'][-1] - lbm.sims_vars['delta_p'][-2])
    return d / max(abs(lbm.sims_vars['delta_p'][-1]), 1e-300) < 1e-10
