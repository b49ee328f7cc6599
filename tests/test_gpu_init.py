"""The initial-interface kernels on the device (idmx_init_*) against the
oracle's restatements of the reference kernels, called directly and through
the drop-in classes (which then set the populations to f_eq(n, 0))."""
import ctypes as C

import numpy as np
import pytest
import sympy as sp

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _device(tenet, fn, dims, *args):
    from idmx import _lib
    from idmx.IdpyCode import IdpyMemory
    V = int(np.prod(dims))
    n = IdpyMemory.Zeros(V, np.float64, tenet)
    u = IdpyMemory.OnDevice(np.full(V * len(dims), 7.0), tenet)
    d = (C.c_int * 3)(*(list(dims) + [1] * (3 - len(dims))))
    _lib.call(fn, tenet.handle, n.ptr, u.ptr, len(dims), d, *args)
    tenet.Sync()
    out = n.D2H(), u.D2H()
    n.Free()
    u.Free()
    return out


@pytest.mark.parametrize('dims', [(17, 12), (9, 10, 7), (64, 64)])
@pytest.mark.parametrize('W', [1.0, 1.7])
def test_init_radial(tenet, dims, W):
    from oracle import cpuref
    n, u = _device(tenet, 'idmx_init_radial', dims, 0.1, 2.0, 3.3, W, 1)
    n_ref, _ = cpuref.init_radial(dims, 0.1, 2.0, 3.3, W)
    assert rel_err(n, n_ref) <= 1e-14 and not u.any()
    # full_flag = 0: gas inside, n reflected about the mean
    n0, _ = _device(tenet, 'idmx_init_radial', dims, 0.1, 2.0, 3.3, W, 0)
    assert np.abs(n0 - (2.1 - n_ref)).max() <= 1e-14


@pytest.mark.parametrize('dims,d,full', [((20, 11), 0, 1), ((20, 11), 1, 0),
                                         ((8, 9, 13), 2, 1), ((7, 12, 5), 1, 1)])
def test_init_flat_cylinder_single(tenet, dims, d, full):
    from oracle import cpuref
    n, u = _device(tenet, 'idmx_init_flat', dims, 0.1, 2.0, 6.0, d, full)
    assert rel_err(n, cpuref.init_flat(dims, 0.1, 2.0, 6.0, d, full)[0]) <= 1e-14
    assert not u.any()
    n, u = _device(tenet, 'idmx_init_cylinder', dims, 0.1, 2.0, 3.6, d, full)
    assert rel_err(n, cpuref.init_cylinder(dims, 0.1, 2.0, 3.6, d, full)[0]) <= 1e-14
    n, u = _device(tenet, 'idmx_init_single_flat', dims, 0.1, 2.0, d, full)
    assert rel_err(n, cpuref.init_single_flat(dims, 0.1, 2.0, d, full)[0]) <= 1e-14


def test_init_rejects_bad_direction(tenet):
    with pytest.raises(RuntimeError):
        _device(tenet, 'idmx_init_flat', (8, 8), 0.1, 2.0, 4.0, 2, 1)


def _classes():
    from idmx.IdpyCode import HIP_T
    from idmx.LBM import LBM, MultiPhase
    n = sp.Symbol('n')
    kw = dict(lang=HIP_T, dim_sizes=(40, 33), xi_stencil=MultiPhase.XIStencils['D2Q9'],
              f_stencil=MultiPhase.FStencils['D2E4'], psi_sym=1 - sp.exp(-n),
              SC_G=-5.0 / 3.0, tau=1.0)
    return (MultiPhase.ShanChenMultiPhase(psi_code='1. - exp(-ln_0)', **kw),
            LBM.ShanChenMultiPhase(psi_code='1. - exp(-ln)', **kw))


def test_class_initialisers_and_populations():
    """InitRadialInterface of both classes (W = 1 Meta, W of the legacy
    class), InitFlatInterface, InitCylinderInterface and the Meta class's
    InitSingleFlatInterface, through the drop-in surface, against the
    oracle; the populations are f_eq(n, 0) = w n."""
    from oracle import cpuref
    meta, legacy = _classes()
    dims = (40, 33)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    cases = [
        (lambda: meta.InitRadialInterface(0.1, 2.0, 9.5), meta,
         cpuref.init_radial(dims, 0.1, 2.0, 9.5, 1.0)[0]),
        (lambda: legacy.InitRadialInterface(0.1, 2.0, 9.5, 2.0), legacy,
         cpuref.init_radial(dims, 0.1, 2.0, 9.5, 2.0)[0]),
        (lambda: meta.InitFlatInterface(0.1, 2.0, 12.0, 1), meta,
         cpuref.init_flat(dims, 0.1, 2.0, 12.0, 1, True)[0]),
        (lambda: legacy.InitCylinderInterface(0.1, 2.0, 7.0, 0), legacy,
         cpuref.init_cylinder(dims, 0.1, 2.0, 7.0, 0, True)[0]),
        (lambda: meta.InitSingleFlatInterface(0.1, 2.0, 0, False), meta,
         cpuref.init_single_flat(dims, 0.1, 2.0, 0, False)[0]),
    ]
    for init, sim, n_ref in cases:
        init()
        n = sim.GetDensityField().ravel()
        assert rel_err(n, n_ref) <= 1e-14
        pop = sim.sims_idpy_memory['pop'].D2H().reshape(9, -1)
        assert np.abs(pop - w[:, None] * n_ref[None, :]).max() <= 1e-14
        assert all(sim.init_status.values())
    meta.End()
    legacy.End()
