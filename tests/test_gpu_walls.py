"""MainLoopWalls (idpy/LBM/MultiPhase.py:1226-1309) through the drop-in
class on the HIP backend against the oracle: fluid sites only, half-way
bounce-back at solid sites, the solid sites' psi set by GetWalls
(MultiPhase.py:157-161). Bar: <= 1e-12 relative on the fluid sites."""
import numpy as np
import pytest
import sympy as sp

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0


def _geometry(dims, seed):
    """Plates at y = 0 and y = L1 - 1, a solid block, wetting psi on solids."""
    rng = np.random.default_rng(seed)
    w = np.ones(tuple(reversed(dims)), dtype=np.uint8)
    if len(dims) == 2:
        w[0, :] = 0
        w[-1, :] = 0
        w[dims[1] // 3:dims[1] // 3 + 4, 5:11] = 0
    else:
        w[:, 0, :] = 0
        w[:, -1, :] = 0
        w[2:5, 4:8, 3:9] = 0
    psi_w = np.where(w == 0, rng.uniform(0.3, 0.7, size=w.shape), 0.0)
    return w, psi_w


@pytest.mark.parametrize('xi,es,dims,nsteps', [('D2Q9', 'D2E4', (64, 37), 300),
                                               ('D2Q9', 'D2E8', (48, 33), 120),
                                               ('D3Q19', 'D3E4', (21, 18, 13), 60)])
def test_mainloopwalls_matches_oracle(tenet, xi, es, dims, nsteps):
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    walls, psi_w = _geometry(dims, 5)
    n0 = droplet_field(dims, seed=45)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                             f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    sim.GetWalls(walls, psi_w)
    sim.MainLoopWalls([0, nsteps // 2, nsteps])
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    st.psi[:] = psi_w.ravel()
    cpuref.meta_run(st, es, G, 1.0, nsteps=nsteps, walls=walls.ravel())
    fl = walls.ravel() == 1
    Q, V = st.Q, st.V
    pop = sim.sims_idpy_memory['pop'].D2H().reshape(Q, V)
    e_pop = rel_err(pop[:, fl].ravel(), st.pop.reshape(Q, V)[:, fl].ravel())
    psi = sim.sims_idpy_memory['psi'].D2H()
    e_psi = rel_err(psi[fl], st.psi[fl])
    # the solid sites keep GetWalls' psi
    assert np.array_equal(psi[~fl], psi_w.ravel()[~fl])
    # ComputeMomentsMeta with walls: fluid-site n, psi and Guo-shifted u
    sim.ComputeMomentsMeta()
    cpuref.meta_moments(st, es, G, walls=walls.ravel())
    n_d = sim.sims_idpy_memory['n'].D2H()
    u_d = sim.sims_idpy_memory['u'].D2H().reshape(st.D, V)
    e_n = rel_err(n_d[fl], st.n[fl])
    e_u = rel_err(u_d[:, fl].ravel(), st.u.reshape(st.D, V)[:, fl].ravel())
    m0 = float(n0.ravel()[fl].sum())
    m1 = float(pop[:, fl].sum())
    sim.End()
    print('rel err pop', e_pop, 'psi', e_psi, 'n', e_n, 'u', e_u, 'mass drift',
          abs(m1 - m0) / m0, flush=True)
    assert e_pop <= 1e-12 and e_psi <= 1e-12 and e_n <= 1e-12
    assert e_u <= 1e-12
    assert abs(m1 - m0) <= 1e-13 * m0


def test_mainloop_ignores_walls_and_moments_honour_them(tenet):
    """MainLoop (no walls) after GetWalls steps the periodic lattice, as on
    the reference, where only MainLoopWalls binds the walls array."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    dims = (40, 31)
    walls, psi_w = _geometry(dims, 6)
    n0 = droplet_field(dims, seed=46)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    sim.GetWalls(walls, psi_w)
    sim.MainLoop([0, 25])
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    cpuref.meta_run(st, 'D2E4', G, 1.0, nsteps=25)
    assert rel_err(sim.sims_idpy_memory['pop'].D2H().ravel(), st.pop) <= 1e-12
    sim.End()


@pytest.mark.parametrize('xi,es,dims,g,nsteps', [('D2Q9', 'D2E4', (64, 37), 2e-5, 250),
                                                 ('D2Q9', 'D2E4', (40, 29), 0.0, 80),
                                                 ('D3Q19', 'D3E4', (21, 18, 13), 1e-5, 50)])
def test_simple_walls_gravity_loop_matches_oracle(tenet, xi, es, dims, g, nsteps):
    """MainLoopSimpleWallsGravity / MainLoopSimpleWalls against the oracle's
    legacy-kernel restatement (pull stream + K_HalfWayBounceBack, F[1] += g):
    fluid-site populations, n, psi and the Guo-shifted u the loop writes."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    walls, psi_w = _geometry(dims, 7)
    n0 = droplet_field(dims, seed=47)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                             f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    sim.GetWalls(walls, psi_w)
    if g:
        sim.MainLoopSimpleWallsGravity([0, nsteps // 3, nsteps], g=g)
    else:
        sim.MainLoopSimpleWalls([0, nsteps])
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    st.psi[:] = psi_w.ravel()
    cpuref.legacy_walls_run(st, es, G, 1.0, walls.ravel(), g, nsteps=nsteps)
    fl = walls.ravel() == 1
    Q, V, D = st.Q, st.V, st.D
    m = sim.sims_idpy_memory
    pop = m['pop'].D2H().reshape(Q, V)
    e_pop = rel_err(pop[:, fl].ravel(), st.pop.reshape(Q, V)[:, fl].ravel())
    e_n = rel_err(m['n'].D2H()[fl], st.n[fl])
    e_psi = rel_err(m['psi'].D2H()[fl], st.psi[fl])
    e_u = rel_err(m['u'].D2H().reshape(D, V)[:, fl].ravel(),
                  st.u.reshape(D, V)[:, fl].ravel())
    # the body force is cleared after the loop: MainLoop is the plain step
    sim.End()
    print('rel err pop', e_pop, 'n', e_n, 'psi', e_psi, 'u', e_u, flush=True)
    assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12
    assert e_u <= 1e-12


@pytest.mark.parametrize('xi,es,dims,walled,nsteps', [('D2Q9', 'D2E4', (64, 37), False, 200),
                                                      ('D2Q9', 'D2E4', (64, 37), True, 200),
                                                      ('D3Q19', 'D3E4', (21, 18, 13), True, 40)])
def test_sc_forcing_loops_match_oracle(tenet, xi, es, dims, walled, nsteps):
    """MainLoopSimpleSC / MainLoopSimpleWallsSC (BGK with the Shan-Chen
    forcing) against the oracle: fluid-site populations, n, psi and the
    velocity the loops write, (u + u_post) / 2."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    walls, psi_w = _geometry(dims, 9)
    n0 = droplet_field(dims, seed=48)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                             f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    if walled:
        sim.GetWalls(walls, psi_w)
        sim.MainLoopSimpleWallsSC([0, nsteps // 2, nsteps])
        st.psi[:] = psi_w.ravel()
        cpuref.legacy_walls_run(st, es, G, 1.0, walls.ravel(), forcing='sc', nsteps=nsteps)
        fl = walls.ravel() == 1
    else:
        sim.MainLoopSimpleSC([0, nsteps])
        cpuref.legacy_walls_run(st, es, G, 1.0, None, forcing='sc', nsteps=nsteps)
        fl = np.ones(st.V, dtype=bool)
    Q, V, D = st.Q, st.V, st.D
    m = sim.sims_idpy_memory
    pop = m['pop'].D2H().reshape(Q, V)
    e_pop = rel_err(pop[:, fl].ravel(), st.pop.reshape(Q, V)[:, fl].ravel())
    e_n = rel_err(m['n'].D2H()[fl], st.n[fl])
    e_psi = rel_err(m['psi'].D2H()[fl], st.psi[fl])
    e_u = rel_err(m['u'].D2H().reshape(D, V)[:, fl].ravel(),
                  st.u.reshape(D, V)[:, fl].ravel())
    sim.End()
    print('rel err pop', e_pop, 'n', e_n, 'psi', e_psi, 'u', e_u, flush=True)
    assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12
    assert e_u <= 1e-12


def test_wall_loops_keep_walls_across_moment_callbacks(tenet):
    """A convergence callback that calls ComputeMomentsMeta between
    checkpoints (as CheckUConvergenceSCMP does) must not change the step of
    MainLoopWalls / MainLoopSimpleWallsGravity."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    dims = (48, 31)
    walls, psi_w = _geometry(dims, 10)
    n0 = droplet_field(dims, seed=49)
    n = sp.Symbol('n')
    calls = []

    def cb(lbm):
        lbm.ComputeMomentsMeta()
        calls.append(1)
        return False

    for loop, g in (('MainLoopWalls', 0.0), ('MainLoopSimpleWallsGravity', 3e-5)):
        sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D2Q9'],
                                 f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                                 psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
        sim.SetDensityField(n0)
        sim.GetWalls(walls, psi_w)
        steps = list(range(0, 61, 20))
        if g:
            sim.MainLoopSimpleWallsGravity(steps, convergence_functions=[cb], g=g)
        else:
            sim.MainLoopWalls(steps, convergence_functions=[cb])
        st = cpuref.State('D2Q9', dims)
        cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
        st.psi[:] = psi_w.ravel()
        if g:
            cpuref.legacy_walls_run(st, 'D2E4', G, 1.0, walls.ravel(), g, nsteps=60)
        else:
            cpuref.meta_run(st, 'D2E4', G, 1.0, nsteps=60, walls=walls.ravel())
        fl = walls.ravel() == 1
        pop = sim.sims_idpy_memory['pop'].D2H().reshape(st.Q, st.V)
        err = rel_err(pop[:, fl].ravel(), st.pop.reshape(st.Q, st.V)[:, fl].ravel())
        sim.End()
        print(loop, 'rel err', err, flush=True)
        assert err <= 1e-12
    assert len(calls) == 6


@pytest.mark.parametrize('loop', ['MainLoopWalls', 'MainLoopSimpleWallsGravity'])
def test_wall_loops_fp32_shifted_close_to_fp64_oracle(tenet, loop):
    """The wall loops on FP32-shifted populations (f - w rho0, rho0 = 1)
    against the FP64 oracle on the fluid sites: within 2e-5 relative after
    100 steps (FP32 storage); fluid mass to 1e-5 (FP32 rounding of the
    density and the collision, a random walk over ~1.7e3 fluid sites:
    measured 1.3e-6; the 256^3 periodic FP32 step drifts 4e-7 in 220)."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    dims = (48, 35)
    walls, psi_w = _geometry(dims, 11)
    n0 = droplet_field(dims, seed=50)
    n = sp.Symbol('n')
    g = 2e-5 if loop.endswith('Gravity') else 0.0
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0, fp32_flag=True)
    sim.SetDensityField(n0)
    sim.GetWalls(walls, psi_w)
    if g:
        sim.MainLoopSimpleWallsGravity([0, 100], g=g)
    else:
        sim.MainLoopWalls([0, 100])
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    st.psi[:] = psi_w.ravel()
    if g:
        cpuref.legacy_walls_run(st, 'D2E4', G, 1.0, walls.ravel(), g, nsteps=100)
    else:
        cpuref.meta_run(st, 'D2E4', G, 1.0, nsteps=100, walls=walls.ravel())
    fl = walls.ravel() == 1
    pop = sim.sims_idpy_memory['pop'].D2H().reshape(st.Q, st.V)
    err = rel_err(pop[:, fl].ravel(), st.pop.reshape(st.Q, st.V)[:, fl].ravel())
    m0 = float(n0.ravel()[fl].sum())
    m1 = float(pop[:, fl].sum())
    sim.End()
    print(loop, 'fp32s rel err', err, 'mass drift', abs(m1 - m0) / m0, flush=True)
    assert err <= 2e-5
    assert abs(m1 - m0) <= 1e-5 * m0


@pytest.mark.parametrize('xi,es,dims,nsteps', [('D2Q9', 'D2E4', (64, 37), 200),
                                               ('D3Q19', 'D3E4', (21, 18, 13), 40)])
def test_simple_loop_matches_oracle_legacy_kernels(tenet, xi, es, dims, nsteps):
    """MainLoopSimple (MultiPhase.py:164-255, the legacy kernels on the
    periodic lattice) against the oracle's legacy loop: populations, n, psi
    and the Guo-shifted u it writes."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    n0 = droplet_field(dims, seed=49)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils[xi],
                             f_stencil=FStencils[es], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    sim.MainLoopSimple([0, nsteps // 2, nsteps])
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    cpuref.legacy_walls_run(st, es, G, 1.0, None, nsteps=nsteps)
    m = sim.sims_idpy_memory
    e_pop = rel_err(m['pop'].D2H().ravel(), st.pop)
    e_n = rel_err(m['n'].D2H(), st.n)
    e_psi = rel_err(m['psi'].D2H(), st.psi)
    e_u = rel_err(m['u'].D2H().ravel(), st.u)
    sim.End()
    print('rel err pop', e_pop, 'n', e_n, 'psi', e_psi, 'u', e_u, flush=True)
    assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12
    assert e_u <= 1e-12
