"""A psi_code outside the closed forms runs as a compiled program on the
device (two-pass step), against the oracle with the same expression
compiled by gcc into its density/psi step, as the reference's CTypes
backend compiles it."""
import numpy as np
import pytest
import sympy as sp

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0
CODES = ['(1. - exp(-ln_0))*(1 + 0.02*ln_0)', '1/2*ln_0 + 1. - exp(-ln_0)',
         '(1. - exp(-ln_0)) - 0.01f*tanh(pow(ln_0, 1.5) - 1)']


@pytest.mark.parametrize('code', CODES)
@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (64, 48)),
                                        ('D3Q19', 'D3E4', (26, 21, 19))])
def test_psi_code_program_matches_gcc_oracle(tenet, tmp_path, code, xi, es, dims):
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    n = sp.Symbol('n')
    keep = cpuref.compile_psi_code(code, str(tmp_path))
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                             f_stencil=FStencils[es], psi_code=code, psi_sym=sp.sqrt(n) + n ** 3,
                             SC_G=G, tau=1.0)
    assert sim.psi_kind == 'program'
    n0 = droplet_field(dims, seed=17)
    sim.SetDensityField(n0)
    sim.MainLoop([0, 100, 250])
    sim.ComputeMomentsMeta()
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0, np.zeros(st.V * st.D))
    cpuref.meta_run(st, es, G, 1.0, psi='code', nsteps=250)
    cpuref.meta_moments(st, es, G, psi='code')
    e_pop = rel_err(sim.sims_idpy_memory['pop'].D2H().ravel(), st.pop)
    e_n = rel_err(sim.GetDensityField().ravel(), st.n)
    e_psi = rel_err(sim.sims_idpy_memory['psi'].D2H(), st.psi)
    e_u = rel_err(sim.sims_idpy_memory['u'].D2H(), st.u)
    sim.End()
    del keep
    print(code, xi, 'pop', e_pop, 'n', e_n, 'psi', e_psi, 'u', e_u, flush=True)
    assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12 and e_u <= 1e-12


def test_psi_program_checks(tenet):
    """Malformed programs are refused (stack under- or overflow at the end,
    constants missing, unknown op codes) and the last valid program stays;
    a program runs on the two-pass step, which needs the n and psi
    buffers."""
    import ctypes as C
    from idmx import _lib
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    eng = SCEngine(tenet, 'D2Q9', 'D2E4', (16, 16), 'program', G, 1.0, psi_code='ln_0')
    bad = [(bytes([0, 0]), []), (bytes([2]), []), (bytes([1]), []), (bytes([0, 99]), [])]
    for ops, consts in bad:
        arr = (C.c_double * max(1, len(consts)))(*consts)
        with pytest.raises(_lib.IdmxError):
            _lib.call('idmx_sc_set_psi_program', eng.handle, ops, len(ops), arr, len(consts))
    n = IdpyMemory.OnDevice(np.ones(256), tenet)
    u = IdpyMemory.Zeros(512, np.float64, tenet)
    psi = IdpyMemory.Zeros(256, np.float64, tenet)
    a, b = PopArray(eng, tenet), PopArray(eng, tenet)
    eng.init_pop(a, n, u)
    with pytest.raises(_lib.IdmxError):
        eng.run(a, b, 2)  # the two-pass step needs n and psi
    eng.run(a, b, 2, n=n, psi=psi)
    for x in (n, u, psi, a, b):
        x.Free()
    eng.close()
