"""Device-resident arrays (the IdpyMemory surface,
idpy/IdpyCode/IdpyMemory.py:343-425, 760-889).

IdpyArrayHIP owns one HBM allocation. The host reads and writes it only
through H2D / D2H (Tenet residency: 'the host touches data only via
D2H/H2D'). Sum and Max reduce on the device in float64.
"""
import ctypes as C

import numpy as np

from .. import _lib


class IdpyArrayHIP:
    def __init__(self, shape, dtype, tenet):
        self.tenet = tenet
        self.shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _lib.call('idmx_alloc', tenet.handle, self.nbytes, C.byref(p))
        self.ptr = p
        tenet._live[id(self)] = self

    @property
    def size(self):
        return int(np.prod(self.shape))

    def H2D(self, array):
        a = np.ascontiguousarray(np.asarray(array, dtype=self.dtype).ravel())
        if a.nbytes != self.nbytes:
            raise ValueError('H2D: size mismatch (%d vs %d bytes)' % (a.nbytes, self.nbytes))
        _lib.call('idmx_h2d', self.tenet.handle, self.ptr, a.ctypes.data, self.nbytes)

    def D2H(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.call('idmx_d2h', self.tenet.handle, out.ctypes.data, self.ptr, self.nbytes)
        return out

    def D2HAt(self, indices):
        """The values at the given flat indices, copied element by element
        (for probes such as the Laplace pressure's two sites, without
        copying the whole field)."""
        out = np.empty(len(indices), dtype=self.dtype)
        isz = self.dtype.itemsize
        for k, i in enumerate(indices):
            i = int(i)
            if not 0 <= i < self.size:
                raise IndexError('D2HAt: index %d out of range' % i)
            _lib.call('idmx_d2h', self.tenet.handle, out[k:].ctypes.data,
                      C.c_void_p(self.ptr.value + i * isz), isz)
        return out

    def SetConst(self, const):
        self.H2D(np.full(self.shape, const, dtype=self.dtype))

    def Free(self):
        if self.ptr is not None and self.tenet.handle is not None:
            _lib.call('idmx_free', self.tenet.handle, self.ptr)
        self.ptr = None
        self.tenet._live.pop(id(self), None)

    def __del__(self):
        try:
            self.Free()
        except Exception:
            pass


def Zeros(shape, dtype=np.float64, tenet=None):
    a = IdpyArrayHIP(shape, dtype, tenet)
    _lib.call('idmx_memset', tenet.handle, a.ptr, 0, a.nbytes)
    return a


def OnDevice(array, tenet=None):
    array = np.asarray(array)
    a = IdpyArrayHIP(array.shape, array.dtype, tenet)
    a.H2D(array)
    return a


def Const(shape, dtype=np.float64, const=0, tenet=None):
    a = IdpyArrayHIP(shape, dtype, tenet)
    a.SetConst(const)
    return a


def _need_f64(array, what):
    if array.dtype != np.float64:
        raise TypeError('%s: the device reductions take float64 arrays, got %s'
                        % (what, array.dtype))


def Sum(array):
    """Device sum in float64 (IdpyMemory.Sum, IdpyMemory.py:865-877)."""
    _need_f64(array, 'Sum')
    out = C.c_double()
    _lib.call('idmx_sum_f64', array.tenet.handle, array.ptr, array.size, C.byref(out))
    return out.value


def Max(array):
    """Device max in float64 (IdpyMemory.Max, IdpyMemory.py:879-889); +inf
    if any entry is NaN."""
    _need_f64(array, 'Max')
    out = C.c_double()
    _lib.call('idmx_max_f64', array.tenet.handle, array.ptr, array.size, C.byref(out))
    return out.value


def MaxAbsDiff(a, b):
    """max |a - b| on the device (float64 arrays of the same size)."""
    _need_f64(a, 'MaxAbsDiff')
    _need_f64(b, 'MaxAbsDiff')
    if a.size != b.size:
        raise ValueError('MaxAbsDiff: sizes differ')
    out = C.c_double()
    _lib.call('idmx_max_abs_diff_f64', a.tenet.handle, a.ptr, b.ptr, a.size, C.byref(out))
    return out.value


def CenterOfMass(n, dims):
    """(sum n, centre of mass) of a float64 device field of lattice size
    dims (x fastest), reduced on the device (K_CenterOfMass + Sum,
    idpy/LBM/MultiPhase.py:108-137)."""
    _need_f64(n, 'CenterOfMass')
    L = list(dims) + [1] * (3 - len(dims))
    out = (C.c_double * 4)()
    _lib.call('idmx_center_of_mass_f64', n.tenet.handle, n.ptr, int(L[0]), int(L[1]), int(L[2]),
              out)
    return out[0], tuple(out[1 + d] for d in range(len(dims)))
