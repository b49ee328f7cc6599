"""Long runs of the benchmark physics on reduced lattices, every field against
the oracle at 1e-12: D2Q9 with the E8 force and several droplets over 2000
steps (rows of 256 on the x-shifted storage), and the D3Q19 MRT droplet over
5000 steps on the LDS-staged fused step on the x-shifted storage (64x4
tiles, partial tiles and a short last segment), compared with the
oracle's moment-space evaluation of the same collision. The lattices are
sized so the serial oracle finishes within about a minute."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0


def _fields(tenet, eng, pop):
    from idmx.IdpyCode import IdpyMemory
    n = IdpyMemory.Zeros(eng.V, np.float64, tenet)
    psi = IdpyMemory.Zeros(eng.V, np.float64, tenet)
    u = IdpyMemory.Zeros(eng.V * eng.D, np.float64, tenet)
    eng.moments(pop, n, psi, u)
    out = n.D2H(), psi.D2H(), u.D2H()
    for x in (n, psi, u):
        x.Free()
    return out


def _run(tenet, eng, n0, nsteps, checkpoints=1):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    n = IdpyMemory.OnDevice(n0, tenet)
    u = IdpyMemory.Zeros(eng.V * eng.D, np.float64, tenet)
    pop, swap = PopArray(eng, tenet), PopArray(eng, tenet)
    eng.init_pop(pop, n, u)
    m0, _ = eng.mass_momentum(pop)
    per = nsteps // checkpoints
    for _ in range(checkpoints):
        if eng.run(pop, swap, per):
            pop, swap = swap, pop
    m1, _ = eng.mass_momentum(pop)
    f = eng.pop_to_host_soa(pop).ravel()
    fields = _fields(tenet, eng, pop)
    for x in (n, u, pop, swap):
        x.Free()
    return f, fields, m0, m1


def test_d2q9_e8_droplets_2000_steps(tenet):
    from idmx import configs
    from idmx.LBM._engine import SCEngine
    from oracle import cpuref
    dims = (200, 200)
    n0 = configs.droplets2d(L=200, count=5, rmin=10.0, rmax=28.0)
    eng = SCEngine(tenet, 'D2Q9', 'D2E8', dims, 'exp', G, 1.0)
    eng.set_launch(256, 1, 32, 10)   # the benchmark's launch for this physics
    f, (n, psi, u), m0, m1 = _run(tenet, eng, n0, 2000, checkpoints=4)
    eng.close()
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0, np.zeros(2 * n0.size))
    cpuref.meta_run(st, 'D2E8', G, 1.0, nsteps=2000)
    f_ref = st.pop.copy()
    cpuref.meta_moments(st, 'D2E8', G)
    errs = [rel_err(f, f_ref), rel_err(n, st.n), rel_err(psi, st.psi), rel_err(u, st.u)]
    print('pop, n, psi, u rel err after 2000 steps', errs, 'mass drift', abs(m1 - m0) / m0,
          flush=True)
    assert max(errs) <= 1e-12
    # rounding of 2000 steps over 40 000 sites: measured 1.1e-13
    assert abs(m1 - m0) <= 5e-13 * m0


@pytest.mark.parametrize('rates', ['unit', 'distinct'])
def test_d3q19_mrt_droplet_5000_steps(tenet, rates):
    """S = I (the benchmark's rates) and a diagonal S of distinct rates."""
    from idmx import configs
    from idmx.LBM import hermite
    from idmx.LBM._engine import SCEngine
    from oracle import cpuref
    dims = (24, 20, 18)
    n0 = configs.droplet(dims, 5.5, 2.0)
    s = np.ones(19) if rates == 'unit' else np.random.default_rng(17).uniform(0.7, 1.5, 19)
    eng = SCEngine(tenet, 'D3Q19', 'D3E4', dims, 'exp', G, 1.0,
                   mrt=hermite.mrt_kernel_tables('D3Q19', np.diag(s)))
    eng.set_launch(64, 4, 16, 10)    # the benchmark's kernel
    f, (n, psi, u), m0, m1 = _run(tenet, eng, n0, 5000, checkpoints=5)
    eng.close()
    st = cpuref.State('D3Q19', dims)
    cpuref.init_pop(st, n0, np.zeros(3 * n0.size))
    cpuref.meta_run_mrt_moment(st, 'D3E4', G, s, nsteps=5000)
    f_ref = st.pop.copy()
    cpuref.meta_moments(st, 'D3E4', G)
    errs = [rel_err(f, f_ref), rel_err(n, st.n), rel_err(psi, st.psi), rel_err(u, st.u)]
    print(rates, 'pop, n, psi, u rel err after 5000 steps', errs, 'mass drift',
          abs(m1 - m0) / m0, flush=True)
    assert max(errs) <= 1e-12
    assert abs(m1 - m0) <= 1e-12 * m0
