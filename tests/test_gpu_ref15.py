"""The current class with the coefficients its generated kernels print
(constants='ref15') against the oracle's literal evaluation of the printed
sums (oracle/ref15_terms.py), and its mass drift against the analytic rate
sum_q printed(w_q) - 1 per step."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0


def _weight_drift(xi_name):
    from decimal import Decimal, getcontext
    from oracle.stencils import XI_W
    getcontext().prec = 15
    r = lambda w: Fr(float(+Decimal(w.numerator) / Decimal(w.denominator)))
    return float(sum(r(w) for w in XI_W[xi_name]) - 1)


def _meta(dims, xi, es, constants='ref15'):
    import sympy as sp
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    return ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                              f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                              psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0, constants=constants)


def test_c1_droplet_ref15_meta_class(tenet):
    """128^2 droplet with seeded noise, E4, 10^4 steps through MainLoop with
    the printed constants, per field against the oracle's printed sums; the
    mass falls by 8.5e-16 per step, as the printed weights sum."""
    from idmx import configs
    from oracle import cpuref
    sim = _meta((128, 128), 'D2Q9', 'D2E4')
    n0 = configs.droplet2d_noise()
    sim.SetDensityField(n0)
    st = cpuref.State('D2Q9', (128, 128))
    cpuref.init_pop_ref15(st, n0, np.zeros(2 * 128 * 128))
    e0 = rel_err(sim.sims_idpy_memory['pop'].D2H().ravel(), st.pop)
    m0 = float(np.sum(st.pop, dtype=np.longdouble))
    sim.MainLoop([0, 1000])
    cpuref.meta_run_ref15(st, 'D2E4', G, 1.0, nsteps=1000)
    err3 = rel_err(sim.sims_idpy_memory['pop'].D2H().ravel(), st.pop)
    print('init', e0, 'after 1e3 steps', err3, flush=True)
    sim.MainLoop(list(range(0, 9001, 1000)))
    cpuref.meta_run_ref15(st, 'D2E4', G, 1.0, nsteps=9000)
    f = sim.sims_idpy_memory['pop'].D2H().ravel()
    err4 = rel_err(f, st.pop)
    drift_gpu = float(np.sum(f, dtype=np.longdouble)) / m0 - 1
    drift_ora = float(np.sum(st.pop, dtype=np.longdouble)) / m0 - 1
    sim.ComputeMomentsMeta()
    cpuref.meta_moments(st, 'D2E4', G)
    e_n = rel_err(sim.GetDensityField().ravel(), st.n)
    e_psi = rel_err(sim.sims_idpy_memory['psi'].D2H(), st.psi)
    e_u = rel_err(sim.sims_idpy_memory['u'].D2H(), st.u)
    sim.End()
    want = 10000 * _weight_drift('D2Q9')
    print('after 1e4 steps', err4, 'n', e_n, 'psi', e_psi, 'u', e_u, 'mass drift gpu',
          drift_gpu, 'oracle', drift_ora, 'printed weights', want, flush=True)
    assert e0 <= 1e-15
    assert err3 <= 1e-12 and err4 <= 1e-12
    assert e_n <= 1e-12 and e_psi <= 1e-12 and e_u <= 1e-12
    assert abs(drift_gpu - drift_ora) <= 1e-14
    assert abs(drift_gpu - want) <= 0.05 * abs(want)


def test_d3q19_ref15_engine_matches_oracle(tenet):
    """D3Q19 E4 droplet on a box with partial tiles, 600 steps of the printed
    step at the default launch and the autotuner's pick."""
    from idmx import tune
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM import ref15
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    from oracle import cpuref
    dims = (40, 36, 32)
    n0 = droplet_field(dims, 11)
    V = int(np.prod(dims))
    st = cpuref.State('D3Q19', dims)
    cpuref.init_pop_ref15(st, n0, np.zeros(3 * V))
    cpuref.meta_run_ref15(st, 'D3E4', G, 1.0, nsteps=600)
    for tuned in (False, True):
        eng = SCEngine(tenet, 'D3Q19', 'D3E4', dims, 'exp', G, 1.0,
                       mrt=ref15.bgk_tables('D3Q19', 1.0), eq_init=ref15.init_table('D3Q19'))
        if tuned:
            best, _ = tune.tune_sim(eng, steps=3, use_cache=False)
            print('launch', best, flush=True)
        n = IdpyMemory.OnDevice(n0, tenet)
        u = IdpyMemory.Zeros(3 * V, np.float64, tenet)
        pop, swap = PopArray(eng, tenet), PopArray(eng, tenet)
        eng.init_pop(pop, n, u)
        if eng.run(pop, swap, 600):
            pop, swap = swap, pop
        f = eng.pop_to_host_soa(pop).ravel()
        for x in (n, u, pop, swap):
            x.Free()
        eng.close()
        err = rel_err(f, st.pop)
        print('tuned' if tuned else 'default', 'rel err', err, flush=True)
        assert err <= 1e-12


def test_mainloopmrt_ref15_matches_dense_oracle(tenet):
    """MainLoopMRT with the printed constants: the dense collision with the
    15-digit tables, against the oracle's dense step with the same tables
    (the literal MRT sums are not restated: parity with the reference's
    printed MRT code is not pinned)."""
    import sympy as sp
    from idmx.LBM import hermite, ref15
    from idmx.LBM.Collisions import RelaxationMatrix
    from idmx.LBM.Equilibria import HermiteEquilibria
    from idmx.LBM.MultiPhase import XIStencils
    from oracle import cpuref
    dims = (34, 30)
    n0 = droplet_field(dims, 5)
    sim = _meta(dims, 'D2Q9', 'D2E4')
    sim.SetDensityField(n0)
    rm = RelaxationMatrix(HermiteEquilibria(xi_stencil=XIStencils['D2Q9']))
    om = sp.symbols('omega_0:9')
    S = rm.GetDiagonalMatrix(list(om))
    rates = np.random.default_rng(4).uniform(0.8, 1.4, 9)
    vals = [(o, sp.Float(v, 17)) for o, v in zip(om, rates)]
    sim.MainLoopMRT([0, 300], relaxation_matrix=S, omega_syms_vals=vals)
    f = sim.sims_idpy_memory['pop'].D2H().ravel()
    sim.End()
    Sn = np.array(S.subs(vals).evalf(), dtype=np.float64)
    A, beta, gamma = ref15.mrt_tables_printed('D2Q9', hermite.mrt_tables('D2Q9', Sn))
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop_ref15(st, n0, np.zeros(2 * st.V))
    cpuref.meta_run_dense_tables(st, 'D2E4', G, (A, beta, gamma), nsteps=300)
    err = rel_err(f, st.pop)
    print('rel err', err, flush=True)
    assert err <= 1e-12
