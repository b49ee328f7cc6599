"""Shared machinery of the two ShanChenMultiPhase classes.

RootLB mirrors idpy/LBM/LBM.py:464-667 (dimensions, strides, constants,
ordering, field getters); _ShanChenCommon holds what both simulation classes
do the same way: kwargs, device memory, initial fields, the MainLoop driver
and the Laplace-pressure analysis. All lattice arithmetic goes through
SCEngine into libidmx.so.
"""
import re
from collections import defaultdict
from functools import reduce

import numpy as np
import sympy as sp

from .. import _lib
from ..IdpyCode import GetParamsClean, GetTenet, HIP_T
from ..IdpyCode import IdpyMemory
from ..IdpyCode import IdpySims as IdpySimsIO
from ..IdpyCode.IdpyMemory import IdpyArrayHIP
from . import hermite, ref15
from .psi_code import compile_psi
from ._engine import SCEngine
from .stencils import estencil_name, xistencil_name


class CustomTypes:
    """Name -> C type table (idpy/Utils/CustomTypes.py:32-53)."""

    def __init__(self, types_dict=None):
        self.types = dict(types_dict or {})

    def Push(self):
        return dict(self.types)

    def __getitem__(self, k):
        return self.types[k]

    def __contains__(self, k):
        return k in self.types


LBMTypes = CustomTypes({'PopType': 'double', 'NType': 'double', 'UType': 'double',
                        'PsiType': 'double', 'ChiType': 'double', 'ThetaType': 'double',
                        'MuType': 'double', 'SCFType': 'double', 'NoiseType': 'double',
                        'ForceType': 'double', 'SType': 'int', 'WType': 'double',
                        'LengthType': 'double', 'FlagType': 'unsigned char'})

_C2NP = {'double': np.float64, 'float': np.float32, 'int': np.int32,
         'unsigned int': np.uint32, 'unsigned char': np.uint8}


def AllTrue(lst):
    return reduce(lambda x, y: x and y, lst)


def OneTrue(lst):
    return reduce(lambda x, y: x or y, lst)


def IndexFromPos(pos, dim_strides):
    index = pos[0]
    for i in range(1, len(pos)):
        index += pos[i] * dim_strides[i - 1]
    return index


def PosFromIndex(index, dim_strides):
    pos = [index % dim_strides[0]]
    pos += [(index // dim_strides[i]) % (dim_strides[i + 1] // dim_strides[i])
            if i < len(dim_strides) - 1 else index // dim_strides[i]
            for i in range(len(dim_strides))]
    return tuple(pos)


class IdpySims:
    """Field dictionaries of a simulation (idpy/IdpyCode/IdpySims.py:36-62)."""

    def __init__(self, *args, **kwargs):
        if not hasattr(self, 'params_dict'):
            self.params_dict = {}
        self.sims_vars = {}
        self.sims_idpy_memory = {}
        self.sims_dump_vars, self.sims_dump_idpy_memory = [], []
        self.sims_not_dump_vars, self.sims_not_dump_idpy_memory = [], []
        self.sims_dump_vars_flag, self.sims_dump_idpy_memory_flag = True, True
        self.aux_idpy_memory, self.aux_vars = [], []

    # -- snapshots (idpy/IdpyCode/IdpySims.py:73-176) ---------------------------
    def DumpSnapshot(self, file_name=None, custom_types=None):
        """<Class>/{vars, idpy_memory, custom_types} into file_name: HDF5 when
        h5py is importable, otherwise the same keys in an npz container."""
        IdpySimsIO.dump(self, file_name, custom_types)

    def ReadSnapshotData(self, file_name=None, full_key=None):
        return IdpySimsIO.read(self, file_name, full_key)

    def CheckSnapshotData(self, file_name=None, full_key=None):
        return IdpySimsIO.check(self, file_name, full_key)

    def ReadSnapshotFull(self):
        pass

    def CleanAuxilliary(self):
        for key in self.aux_vars:
            self.sims_vars.pop(key, None)
        for key in self.aux_idpy_memory:
            a = self.sims_idpy_memory.pop(key, None)
            if a is not None:
                a.Free()
        self.aux_vars, self.aux_idpy_memory = [], []


class PopArray(IdpyArrayHIP):
    """Population buffer. The device layout is SoA; D2H/H2D present the
    ordering the simulation was created with ('gpu': pop[g + q*V], 'cpu':
    pop[q + g*Q], idpy/LBM/LBM.py:529-537) in float64 (FP32-shifted storage is
    converted)."""

    def __init__(self, engine, tenet, aos=False):
        self.engine, self.aos = engine, aos
        IdpyArrayHIP.__init__(self, engine.Q * engine.V, engine.pop_dtype, tenet)

    def D2H(self):
        f = self.engine.pop_to_host_soa(self)
        return (f.T if self.aos else f).ravel().copy()

    def H2D(self, array):
        a = np.asarray(array, dtype=np.float64).ravel()
        e = self.engine
        f = a.reshape(e.V, e.Q).T if self.aos else a.reshape(e.Q, e.V)
        e.pop_from_host_soa(self, f)

    def RawD2H(self):
        return IdpyArrayHIP.D2H(self)


class VelArray(IdpyArrayHIP):
    """Velocity buffer, SoA on the device; 'cpu' ordering is u[d + g*DIM]."""

    def __init__(self, V, D, tenet, aos=False):
        self.V, self.Dd, self.aos = V, D, aos
        IdpyArrayHIP.__init__(self, V * D, np.float64, tenet)

    def D2H(self):
        u = IdpyArrayHIP.D2H(self).reshape(self.Dd, self.V)
        return (u.T if self.aos else u).ravel().copy()

    def H2D(self, array):
        a = np.asarray(array, dtype=np.float64).ravel()
        u = a.reshape(self.V, self.Dd).T if self.aos else a.reshape(self.Dd, self.V)
        IdpyArrayHIP.H2D(self, np.ascontiguousarray(u))


def psi_kind_from(psi_sym, psi_code):
    """Maps the user's pseudopotential onto a kernel form. psi_sym (sympy in
    the symbol n) decides; psi_code (a C expression in ln / ln_0) is the
    fallback. Unknown forms raise."""
    n = sp.Symbol('n')
    cands = {'exp': 1 - sp.exp(-n), 'expinv': sp.exp(-1 / n), 'linear': n}
    if psi_sym is not None:
        expr = sp.sympify(psi_sym)
        syms = list(expr.free_symbols)
        if len(syms) == 1 and syms[0] != n:
            expr = expr.subs(syms[0], n)
        for k, c in cands.items():
            if sp.simplify(expr - c) == 0:
                return k
    if psi_code is not None:
        s = re.sub(r'\((PsiType|NType|double)\)', '', psi_code)
        s = re.sub(r'\s+', '', s).replace('ln_0', 'N').replace('ln', 'N')
        s = s.replace('1.0', '1').replace('1.', '1')
        table = {'1-exp(-N)': 'exp', 'exp(-1/N)': 'expinv', 'N': 'linear'}
        if s in table:
            return table[s]
        # any other C expression: compiled into a stack program that the
        # kernels evaluate (raises NotImplementedError outside the grammar)
        compile_psi(psi_code)
        return 'program'
    raise NotImplementedError(
        'pseudopotential %r without psi_code is not one of 1 - exp(-n), exp(-1/n), n'
        % (psi_sym,))


RootLB_Dump_def_name = 'RootLB_dump.hdf5'


class RootLB(IdpySims):
    """Dimensions, strides, constants and field getters
    (idpy/LBM/LBM.py:464-667)."""

    def __init__(self, *args, **kwargs):
        if not hasattr(self, 'params_dict'):
            self.params_dict = {}
        self.kwargs = GetParamsClean(kwargs, [self.params_dict],
                                     needed_params=['custom_types', 'dim_sizes', 'xi_stencil',
                                                    'root_dim_sizes', 'root_strides',
                                                    'root_coord'])
        if 'dim_sizes' not in self.params_dict:
            raise Exception("Missing parameter 'dim_sizes'")
        if 'xi_stencil' not in self.params_dict:
            raise Exception("Missing parameter 'xi_stencil'")
        for k, v in (('root_dim_sizes', 'L'), ('root_strides', 'STR'), ('root_coord', 'x')):
            self.params_dict.setdefault(k, v)
        IdpySims.__init__(self, *args, **self.kwargs)

        dim_sizes = self.params_dict['dim_sizes']
        xi_stencil = self.params_dict['xi_stencil']
        D = len(dim_sizes)
        sizes = np.array(dim_sizes, dtype=np.int32)
        strides = np.array([int(np.prod(sizes[:i + 1])) for i in range(D - 1)], dtype=np.int32)
        self.sims_vars['DIM'] = D
        self.sims_vars['dim_sizes'] = sizes
        self.sims_vars['dim_strides'] = strides if len(strides) else sizes
        self.sims_vars['dim_center'] = np.array([s // 2 for s in sizes], dtype=np.int32)
        self.sims_vars['V'] = int(np.prod(sizes))
        self.sims_vars['Q'] = len(xi_stencil['XIs'])
        self.sims_vars['XI_list'] = np.array([x for xi in xi_stencil['XIs'] for x in xi],
                                             dtype=np.int32)
        self.sims_vars['W_list'] = np.array([float(w) for w in xi_stencil['Ws']])
        self.sims_vars['c2'] = float(xi_stencil['c2'])

        ol = defaultdict(lambda: defaultdict(dict))
        ol['gpu']['pop'] = lambda pos, q: str(pos) + ' + ' + str(q) + ' * V'
        ol['cpu']['pop'] = lambda pos, q: str(q) + ' + ' + str(pos) + ' * Q'
        ol['gpu']['u'] = lambda pos, d: str(pos) + ' + ' + str(d) + ' * V'
        ol['cpu']['u'] = lambda pos, d: str(d) + ' + ' + str(pos) + ' * DIM'
        self.sims_vars['ordering_lambdas'] = ol

        c2 = self.sims_vars['c2']
        self.constants = {'V': self.sims_vars['V'], 'Q': self.sims_vars['Q'], 'DIM': D,
                          'CM2': 1 / c2, 'CM4': 1 / (c2 ** 2), 'PI': np.pi}
        for i, s in enumerate(sizes):
            self.constants[self.params_dict['root_dim_sizes'] + '_' + str(i)] = int(s)
        for i, s in enumerate(self.sims_vars['dim_strides']):
            self.constants[self.params_dict['root_strides'] + '_' + str(i)] = int(s)
        self.sims_idpy_memory = {'pop': None, 'dim_sizes': None, 'dim_strides': None,
                                 'dim_center': None, 'XI_list': None, 'W_list': None}

    def GetDensityField(self):
        n = self.sims_idpy_memory['n'].D2H()
        return n.reshape(np.flip(self.sims_vars['dim_sizes']))

    def GetVelocityField(self):
        if 'set_ordering' not in self.params_dict:
            raise Exception("Variable 'set_ordering' not present in 'params_dict'")
        u = self.sims_idpy_memory['u'].D2H()
        D, sizes = self.sims_vars['DIM'], self.sims_vars['dim_sizes']
        if self.params_dict['set_ordering'] == 'cpu':
            dims = np.flip(np.append([D], sizes))
        else:
            dims = np.flip(np.append(sizes, [D]))
        return u.reshape(dims)

    def DumpPopSnapshot(self, file_name=None):
        """DumpSnapshot with the simulation's custom_types
        (idpy/LBM/LBM.py:605-607)."""
        self.DumpSnapshot(file_name=RootLB_Dump_def_name if file_name is None else file_name,
                          custom_types=self.custom_types)

    def GetPopulations(self):
        if 'set_ordering' not in self.params_dict:
            raise Exception("Variable 'set_ordering' not present in 'params_dict'")
        pop = self.sims_idpy_memory['pop'].D2H()
        Q, sizes = self.sims_vars['Q'], self.sims_vars['dim_sizes']
        if self.params_dict['set_ordering'] == 'cpu':
            dims = np.flip(np.append([Q], sizes))
        else:
            dims = np.flip(np.append(sizes, [Q]))
        return pop.reshape(dims)


class _ShanChenCommon(RootLB):
    """What both ShanChenMultiPhase classes share."""

    # constructor kwargs of the HIP backend on top of the reference's
    _EXTRA_PARAMS = ['precision', 'rho0', 'step_mode', 'constants']
    # whether the reference class prints its coefficients (constants='ref15')
    _PRINTED_CONSTANTS = False

    def _setup(self, needed, args, kwargs, psi_var):
        if not hasattr(self, 'params_dict'):
            self.params_dict = {}
        self.kwargs = GetParamsClean(kwargs, [self.params_dict],
                                     needed_params=needed + self._EXTRA_PARAMS)
        for k, msg in (('f_stencil', "Missing 'f_stencil'"), ('tau', "Missing 'tau'"),
                       ('SC_G', "Missing 'SC_G'"),
                       ('psi_code', "Missing 'psi_code', e.g. psi_code = '1. - exp(-%s)'" % psi_var),
                       ('lang', 'Param lang = HIP_T is needed'),
                       ('psi_sym', "Missing sympy expression for the pseudo-potential, "
                                   "parameter 'psi_sym'")):
            if k not in self.params_dict:
                raise Exception(msg)
        self.optimizer_flag = self.params_dict.get('optimizer_flag', True)
        self.params_dict.setdefault('empty_sim', False)
        self.params_dict.setdefault('precision',
                                    'fp32s' if self.params_dict.get('fp32_flag') else 'fp64')
        self.params_dict.setdefault('step_mode', 'fused')
        # 'exact': full-double constants (mass conserved); 'ref15': the 15-digit
        # coefficients the current class prints into its generated kernels
        self.params_dict.setdefault('constants', 'exact')
        if self.params_dict['constants'] not in ('exact', 'ref15'):
            raise ValueError("constants must be 'exact' or 'ref15'")
        if self.params_dict['constants'] == 'ref15':
            if not self._PRINTED_CONSTANTS:
                raise ValueError("constants='ref15' belongs to the current class "
                                 "(idmx.LBM.MultiPhase.ShanChenMultiPhase); this class reads "
                                 "full-double constants")
            if self.params_dict['precision'] != 'fp64':
                raise ValueError("constants='ref15' needs FP64 populations")
        self.tenet = GetTenet(self.params_dict)
        self.custom_types = self.params_dict.get('custom_types', LBMTypes)
        if self.params_dict['precision'] == 'fp32s':
            t = self.custom_types.Push()
            t.update({'PopType': 'float'})
            self.custom_types = CustomTypes(t)
        RootLB.__init__(self, *args, custom_types=self.custom_types, **self.kwargs)

        self.xi_name = xistencil_name(self.params_dict['xi_stencil'])
        self.e_name = estencil_name(self.params_dict['f_stencil'])
        if self.xi_name is None:
            raise NotImplementedError('xi_stencil must be D2Q9 or D3Q19')
        if self.e_name is None or int(self.e_name[1]) != self.sims_vars['DIM']:
            raise NotImplementedError('f_stencil must be one of D%dE4/E6/E8'
                                      % self.sims_vars['DIM'])
        self.psi_kind = psi_kind_from(self.params_dict['psi_sym'], self.params_dict['psi_code'])

    def _init_vars(self):
        f = self.params_dict['f_stencil']
        es = f['Es'] if 'Es' in f else f['XIs']
        self.sims_vars['tau'] = self.params_dict['tau']
        self.sims_vars['QE'] = len(es)
        self.sims_vars['E_list'] = np.array([x for e in es for x in e], dtype=np.int32)
        self.sims_vars['EW_list'] = np.array([float(w) for w in f['Ws']])
        self.sims_vars['SC_G'] = self.params_dict['SC_G']
        self.sims_vars['e2_val'] = self.params_dict.get('e2_val', 1)
        self.sims_vars['psi_sym'] = self.params_dict['psi_sym']
        self.sims_vars['n_sym'] = sp.symbols('n')
        self.constants['QE'] = self.sims_vars['QE']
        self.constants['SC_G'] = self.params_dict['SC_G']
        self.constants['OMEGA'] = 1. / self.params_dict['tau']
        bs = self.params_dict.get('block_size', 128)
        self.sims_vars['grid'] = ((self.sims_vars['V'] + bs - 1) // bs, 1, 1)
        self.sims_vars['block'] = (bs, 1, 1)

    def _make_engine(self, mrt=None, forcing='guo', iso_filter=False):
        rho0 = self.params_dict.get('rho0', 1.0)
        eq_init = None
        if self.params_dict['constants'] == 'ref15' and forcing == 'guo' and not iso_filter:
            # the printed coefficients, through the dense collision tables
            if mrt is None:
                mrt = ref15.bgk_tables(self.xi_name, self.params_dict['tau'])
            else:
                mrt = ref15.mrt_tables_printed(self.xi_name, mrt)
            eq_init = ref15.init_table(self.xi_name)
        return SCEngine(self.tenet, self.xi_name, self.e_name,
                        tuple(int(s) for s in self.sims_vars['dim_sizes']), self.psi_kind,
                        self.params_dict['SC_G'], self.params_dict['tau'],
                        precision=self.params_dict['precision'], rho0=rho0, mrt=mrt,
                        forcing=forcing, iso_filter=iso_filter, eq_init=eq_init,
                        psi_code=self.params_dict['psi_code']
                        if self.psi_kind == 'program' else None)

    def _scf_engine(self):
        """Engine with the Shan-Chen forcing (BGK), created on first use."""
        if getattr(self, '_scf', None) is None:
            self._scf = self._make_engine(forcing='sc')
        return self._scf

    def _iso_engine(self):
        """Engine of BGK followed by the isotropy filter, created on first use."""
        if getattr(self, '_iso', None) is None:
            self._iso = self._make_engine(iso_filter=True)
        return self._iso

    def _init_memory(self, pop_aos, u_aos):
        t = self.tenet
        V, D = self.sims_vars['V'], self.sims_vars['DIM']
        self.engine = self._make_engine()
        self._mrt_engines = {}
        m = self.sims_idpy_memory
        m['pop'] = PopArray(self.engine, t, aos=pop_aos)
        m['pop_swap'] = PopArray(self.engine, t, aos=pop_aos)
        m['n'] = IdpyMemory.Zeros(V, np.float64, t)
        m['psi'] = IdpyMemory.Zeros(V, np.float64, t)
        m['u'] = VelArray(V, D, t, aos=u_aos)
        IdpyMemory.IdpyArrayHIP.H2D(m['u'], np.zeros(V * D))
        m['dim_sizes'] = IdpyMemory.OnDevice(self.sims_vars['dim_sizes'], t)
        m['dim_strides'] = IdpyMemory.OnDevice(self.sims_vars['dim_strides'], t)
        m['dim_center'] = IdpyMemory.OnDevice(self.sims_vars['dim_center'], t)
        m['XI_list'] = IdpyMemory.OnDevice(self.sims_vars['XI_list'], t)
        m['W_list'] = IdpyMemory.OnDevice(self.sims_vars['W_list'], t)
        m['E_list'] = IdpyMemory.OnDevice(self.sims_vars['E_list'], t)
        m['EW_list'] = IdpyMemory.OnDevice(self.sims_vars['EW_list'], t)

    # -- initial fields (host-side geometry, then H2D) -----------------------
    def _set_n_u0(self, n):
        m = self.sims_idpy_memory
        m['n'].H2D(np.asarray(n, dtype=np.float64))
        IdpyMemory.IdpyArrayHIP.H2D(m['u'], np.zeros(self.sims_vars['V'] * self.sims_vars['DIM']))
        self.init_status['n'] = True
        self.init_status['u'] = True

    def InitPopulations(self):
        """K_InitPopulations: pop = f_eq(n, u) on the device."""
        if not AllTrue([self.init_status['n'], self.init_status['u']]):
            raise Exception('Fields u and n are not initialized')
        m = self.sims_idpy_memory
        self.engine.init_pop(m['pop'], m['n'], m['u'])
        self.init_status['pop'] = True

    def _device_init(self, fn, *args):
        """Runs one of the initial-interface kernels (idmx_init_*) on n and u,
        then marks both initialised."""
        import ctypes as C
        m = self.sims_idpy_memory
        sizes = [int(v) for v in self.sims_vars['dim_sizes']]
        dims = (C.c_int * 3)(*(sizes + [1] * (3 - len(sizes))))
        _lib.call(fn, self.tenet.handle, m['n'].ptr, m['u'].ptr, len(sizes), dims, *args)
        self.tenet.Sync()
        self.init_status['n'] = True
        self.init_status['u'] = True

    def InitFlatInterface(self, n_g, n_l, width, direction=0, full_flag=True):
        """K_InitFlatInterface (idpy/LBM/LBM.py:1745-1784) with
        F_NFlatProfilePeriodic (idpy/LBM/MultiPhaseKernels.py:55-74), on the
        device (idmx_init_flat)."""
        self.sims_vars.update({'init_type': 'flat', 'n_g': n_g, 'n_l': n_l,
                               'full_flag': full_flag, 'width': width, 'direction': direction})
        self._device_init('idmx_init_flat', float(n_g), float(n_l), float(width),
                          int(direction), int(bool(full_flag)))
        self.InitPopulations()

    def InitCylinderInterface(self, n_g, n_l, R, direction=0, full_flag=True):
        """K_InitCylinderInterface (idpy/LBM/LBM.py:1828-1874), on the device
        (idmx_init_cylinder)."""
        self.sims_vars.update({'init_type': 'cylinder', 'n_g': n_g, 'n_l': n_l,
                               'full_flag': full_flag, 'R': R, 'direction': direction})
        self._device_init('idmx_init_cylinder', float(n_g), float(n_l), float(R),
                          int(direction), int(bool(full_flag)))
        self.InitPopulations()

    def _init_radial(self, n_g, n_l, R, W, full_flag):
        """K_InitRadialInterface on the device (idmx_init_radial)."""
        self.sims_vars.update({'init_type': 'radial', 'n_g': n_g, 'n_l': n_l,
                               'full_flag': full_flag, 'R': R})
        self._device_init('idmx_init_radial', float(n_g), float(n_l), float(R), float(W),
                          int(bool(full_flag)))
        self.InitPopulations()

    def SetDensityField(self, n):
        """Sets n (array in (z,) y, x order or flat lexicographic), u = 0 and
        the populations to f_eq(n, 0). The configurations with seeded noise or
        many droplets are set this way, as on the reference (H2D of n, then
        InitPopulations)."""
        self._set_n_u0(np.asarray(n, dtype=np.float64).ravel())
        self.InitPopulations()

    # -- the time loop ------------------------------------------------------
    def _run_loop(self, time_steps, convergence_functions, profiling, engine, fields_u):
        """MainLoop driver (idpy/LBM/MultiPhase.py:1015-1097): runs each
        checkpoint interval on the device, then the convergence callbacks."""
        _all_init = [self.init_status[k] for k in self.init_status]
        if not AllTrue(_all_init):
            print(self.init_status)
            raise Exception('Hydrodynamic variables/populations not initialized')
        m = self.sims_idpy_memory
        two_pass = self.params_dict.get('step_mode', 'fused') == 'two_pass'
        old_step, prof = 0, {}
        for step in time_steps[1:]:
            print('Step:', step)
            nst = step - old_step
            if nst > 0:
                step_ms = np.zeros(2 * nst, dtype=np.float32) if profiling else None
                swapped = engine.run(m['pop'], m['pop_swap'], nst, n=m['n'], psi=m['psi'],
                                     u=fields_u, write_fields=True, two_pass=two_pass,
                                     step_ms=step_ms)
                if swapped:
                    m['pop'], m['pop_swap'] = m['pop_swap'], m['pop']
                if profiling:
                    # IdpyLoopProfile's layout: {sequence: {kernel: seconds per
                    # step, 'device_name': ...}} (idpy/IdpyCode/IdpyCode.py:
                    # 1913-2085); walls, a body force and the Shan-Chen or
                    # filtered collisions run the two-pass step
                    k0, k1 = engine.step_kernel_names(two_pass)
                    t = step_ms.reshape(nst, 2).astype(np.float64) / 1e3
                    seq = {k1: t[:, 1], 'device_name': self.tenet.GetDeviceName()}
                    if k0 is not None:
                        seq[k0] = t[:, 0]
                    prof[step] = {0: seq}
            old_step = step
            if len(convergence_functions):
                checks = [c_f(self) for c_f in convergence_functions]
                if OneTrue(checks):
                    break
        return prof if profiling else None

    def MassMomentum(self):
        """FP64 device reduction of total mass and momentum of pop."""
        return self.engine.mass_momentum(self.sims_idpy_memory['pop'])

    # -- analysis -----------------------------------------------------------
    def ComputeCenterOfMass(self):
        """(mass, centre of mass) of n, reduced on the device in FP64
        (idpy/LBM/MultiPhase.py:108-137)."""
        return IdpyMemory.CenterOfMass(self.sims_idpy_memory['n'],
                                       self.sims_vars['dim_sizes'])

    def DeltaPLaplace(self, inside=None, outside=None):
        """Laplace pressure p(centre) - p(V-1) (idpy/LBM/MultiPhase.py:1690-1756)."""
        if inside is None:
            _mass, _inside = self.ComputeCenterOfMass()
            self.sims_vars['cm_coords'] = np.array(_inside)
            self.sims_vars['mass'] = _mass
            _chk = any(abs(self.sims_vars['dim_center'][d] - _inside[d]) > 1e-3
                       for d in range(self.sims_vars['DIM']))
            self.sims_vars['is_centered'] = not _chk
            if _chk:
                print('Center of mass: ', _inside, '; Center of the system: ',
                      self.sims_vars['dim_center'])
            inside = IndexFromPos(self.sims_vars['dim_center'].tolist(),
                                  self.sims_vars['dim_strides'])
        if outside is None:
            outside = self.sims_vars['V'] - 1
        nio = self.sims_idpy_memory['n'].D2HAt([inside, outside])
        self.sims_vars['n_in_n_out'] = nio
        p_in, p_out = self.PBulk(nio[0]), self.PBulk(nio[1])
        return p_in, p_out, p_in - p_out

    def SurfaceTensionLaplace(self, n_g=None, n_l=None):
        """Surface tension of a centred droplet from Laplace's law with the
        equimolar radius, from FP64 device reductions of n (call
        ComputeMomentsMeta first). n_l / n_g default to n at the centre and
        at site V-1 (the DeltaPLaplace probes). The equimolar radius R
        satisfies M = V n_g + (n_l - n_g) |B_R| with |B_R| = pi R^2 (2D) or
        4 pi R^3 / 3 (3D); sigma = dP R (2D) or dP R / 2 (3D). Returns
        (sigma, R, dP)."""
        _p_in, _p_out, dP = self.DeltaPLaplace()
        nio = self.sims_vars['n_in_n_out']
        n_l = nio[0] if n_l is None else n_l
        n_g = nio[1] if n_g is None else n_g
        mass = self.sims_vars['mass']
        vol = (mass - self.sims_vars['V'] * n_g) / (n_l - n_g)
        if self.sims_vars['DIM'] == 2:
            R = np.sqrt(vol / np.pi)
            sigma = dP * R
        else:
            R = (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)
            sigma = 0.5 * dP * R
        return float(sigma), float(R), float(dP)

    def PBulk(self, n):
        """p = n c2 + G e2 psi(n)^2 / 2 (idpy/LBM/MultiPhase.py:1759-1768)."""
        if 'psi_f' not in self.sims_vars:
            self.sims_vars['psi_f'] = sp.lambdify(self.sims_vars['n_sym'],
                                                  self.sims_vars['psi_sym'])
            self.sims_not_dump_vars += ['psi_f']
        return (n * self.sims_vars['c2'] + 0.5 * self.sims_vars['SC_G']
                * self.sims_vars['e2_val'] * (self.sims_vars['psi_f'](n)) ** 2)

    # -- MRT ----------------------------------------------------------------
    def _mrt_engine(self, relaxation_matrix, omega_syms_vals):
        S = relaxation_matrix
        if isinstance(S, sp.MatrixBase):
            for sym, val in (omega_syms_vals or []):
                S = S.subs(sym, val)
            if S.free_symbols:
                raise Exception('Some of the rates values are missing!')
            S = np.array(S.evalf(), dtype=np.float64)
        S = np.asarray(S, dtype=np.float64)
        key = S.tobytes()
        if key not in self._mrt_engines:
            self._mrt_engines[key] = self._make_engine(
                mrt=hermite.mrt_tables(self.xi_name, S)
                if self.params_dict['constants'] == 'ref15'
                else hermite.mrt_kernel_tables(self.xi_name, S))
        return self._mrt_engines[key]

    def End(self):
        for e in list(getattr(self, '_mrt_engines', {}).values()):
            e.close()
        if getattr(self, '_scf', None) is not None:
            self._scf.close()
        if getattr(self, '_iso', None) is not None:
            self._iso.close()
        if getattr(self, 'engine', None) is not None:
            self.engine.close()
        self.tenet.FreeMemoryDict(memory_dict=self.sims_idpy_memory)
        self.tenet.End()
