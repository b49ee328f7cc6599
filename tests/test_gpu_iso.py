"""MainLoopIsoFilter (idpy/LBM/MultiPhase.py:1099-1223) through the drop-in
class on the HIP backend against the oracle: the Meta step with the
isotropy filter (K_IsotropyFilter, idpy/LBM/LBMKernelsMeta.py:895-1080)
between collision and push stream. Bar: <= 1e-12 relative, mass conserved
to 1e-13."""
import numpy as np
import pytest
import sympy as sp

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0


def _sim(dims, es, **kw):
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    return ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D2Q9'],
                              f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                              psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0, **kw)


@pytest.mark.parametrize('es,dims,nsteps', [('D2E4', (64, 37), 300), ('D2E8', (48, 40), 200)])
def test_iso_filter_loop_matches_oracle(tenet, es, dims, nsteps):
    from oracle import cpuref
    n0 = droplet_field(dims, seed=52)
    sim = _sim(dims, es)
    sim.SetDensityField(n0)
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    m0 = st.pop.sum()
    sim.MainLoopIsoFilter([0, nsteps // 3, nsteps])
    cpuref.iso_run(st, es, G, 1.0, nsteps=nsteps)
    m = sim.sims_idpy_memory
    pop = m['pop'].D2H().ravel()
    e_pop = rel_err(pop, st.pop)
    e_n = rel_err(m['n'].D2H().ravel(), st.n)
    e_psi = rel_err(m['psi'].D2H().ravel(), st.psi)
    drift = abs(pop.sum() - m0) / m0
    # the moments after the loop come from the Guo engine, as on the reference
    sim.ComputeMomentsMeta()
    u = m['u'].D2H()
    sim.End()
    print('rel err pop', e_pop, 'n', e_n, 'psi', e_psi, 'mass drift', drift, flush=True)
    assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12
    assert drift <= 1e-13
    assert np.isfinite(u).all()


def test_iso_filter_loop_differs_from_mainloop_and_rejects_other_setups(tenet):
    """The filter changes the trajectory (it is not a no-op), and the loop
    is refused where the filter is not defined (D3Q19, FP32)."""
    dims = (40, 32)
    n0 = droplet_field(dims, seed=53)
    a, b = _sim(dims, 'D2E4'), _sim(dims, 'D2E4')
    for s in (a, b):
        s.SetDensityField(n0)
    a.MainLoopIsoFilter([0, 50])
    b.MainLoop([0, 50])
    d = rel_err(a.sims_idpy_memory['pop'].D2H(), b.sims_idpy_memory['pop'].D2H())
    a.End()
    b.End()
    assert d > 1e-4
    c = _sim(dims, 'D2E4', fp32_flag=True)
    c.SetDensityField(n0)
    with pytest.raises(NotImplementedError):
        c.MainLoopIsoFilter([0, 5])
    c.End()
    from idmx.LBM._engine import SCEngine
    with pytest.raises(ValueError):
        SCEngine(tenet, 'D3Q19', 'D3E4', (8, 8, 8), 'exp', G, 1.0, iso_filter=True)
