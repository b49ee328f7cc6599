"""Language tags, Tenet (device context) and kwargs plumbing.

Mirrors idpy.IdpyCode (idpy/IdpyCode/__init__.py:115-284): a simulation class
picks its backend through the 'lang' kwarg and GetTenet. This package has a
single backend, HIP on gfx950, selected with lang = HIP_T. The reference's
other tags are accepted by name so that notebooks can be retargeted by
changing one value (the way scripts/smoke_papers.py:113-127 rewrites
preferred_lang); selecting them here raises.
"""
import ctypes as C

from .. import _lib

HIP_T = 'hip'
IDPY_T = 'idpy'
# The reference's tags (idpy/CUDA/__init__.py:41, OpenCL/__init__.py:41,
# CTypes/__init__.py:44, Metal/__init__.py:44).
CUDA_T = 'pycuda'
OCL_T = 'pyopencl'
CTYPES_T = 'ctypes'
METAL_T = 'pymetallic'

idpy_langs_dict = {'HIP_T': HIP_T}
idpy_langs_human_dict = {HIP_T: 'HIP (gfx950)'}
idpy_langs_list = [HIP_T]
idpy_langs_sys = {HIP_T: True, CUDA_T: False, OCL_T: False, CTYPES_T: False, METAL_T: False}


def GetParamsClean(kwargs, _a_params_dict, needed_params=None):
    """Moves the needed_params out of kwargs into _a_params_dict[0]
    (idpy/IdpyCode/__init__.py:274-284)."""
    for var in needed_params:
        if var in kwargs:
            _a_params_dict[0][var] = kwargs[var]
            del kwargs[var]
    return kwargs


class Tenet:
    """One HIP device context with its own stream (the role of idpy's
    per-backend Tenet, e.g. idpy/CTypes/CTypes.py:59-83). Arrays allocated
    through it live in HBM; the host touches them only via H2D/D2H."""

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        _lib.call('idmx_ctx_create', int(device), C.byref(h))
        self.handle = h
        self._live = {}

    def GetKind(self):
        return 'gpu'

    def GetDeviceName(self):
        buf = C.create_string_buffer(256)
        _lib.call('idmx_device_name', self.handle, buf, 256)
        return buf.value.decode()

    def Sync(self):
        _lib.call('idmx_sync', self.handle)

    def FreeMemoryDict(self, memory_dict=None):
        for key in list(memory_dict or {}):
            a = memory_dict[key]
            if hasattr(a, 'Free'):
                a.Free()

    def End(self):
        if self.handle:
            # arrays, simulations and batches still open on this context are
            # released first: their handles are not valid once it is gone
            for a in list(self._live.values()):
                a.close() if hasattr(a, 'close') else a.Free()
            self._live.clear()
            _lib.call('idmx_ctx_destroy', self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.End()
        except Exception:
            pass


def GetTenet(params_dict):
    """A new HIP tenet on params_dict['device'] (default 0), as the
    reference's GetTenet builds a new backend object per simulation
    (idpy/IdpyCode/__init__.py:166-211)."""
    lang = params_dict.get('lang', None)
    if lang != HIP_T:
        raise Exception("lang = %r is not available in this package: use lang = HIP_T" % (lang,))
    t = Tenet(params_dict.get('device', 0))
    print('HIP: ', t.GetDeviceName())
    return t
