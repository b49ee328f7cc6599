"""The drop-in classes, constructed as a notebook does, on the GPU: the
legacy idpy.LBM.LBM.ShanChenMultiPhase surface and the Meta
idpy.LBM.MultiPhase.ShanChenMultiPhase surface, compared with the oracle
(workload: D2Q9 128^2 droplet with seeded noise, E4, G = -5, tau = 1)."""
import numpy as np
import pytest
import sympy as sp

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _legacy(nsteps_list, **kw):
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.LBM import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(128, 128), xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0, **kw)
    return sim


def test_legacy_class_c1_matches_oracle():
    from idmx import configs
    from oracle import cpuref
    n0 = configs.droplet2d_noise()
    sim = _legacy(None)
    sim.SetDensityField(n0)
    sim.MainLoop([0, 100, 300])
    st = cpuref.State('D2Q9', (128, 128))
    cpuref.init_pop(st, n0, np.zeros(2 * 128 * 128))
    cpuref.legacy_run(st, 'D2E4', -5.0 / 3.0, 1.0, nsteps=300)
    assert rel_err(sim.GetDensityField().ravel(), st.n) <= 1e-12
    assert rel_err(sim.sims_idpy_memory['psi'].D2H(), st.psi) <= 1e-12
    e_u = rel_err(sim.sims_idpy_memory['u'].D2H(), st.u)
    print('u rel err', e_u, flush=True)
    assert e_u <= 1e-12
    assert rel_err(sim.sims_idpy_memory['pop'].D2H(), st.pop) <= 1e-12
    sim.End()


@pytest.mark.parametrize('ordering', ['gpu', 'cpu'])
def test_meta_class_orderings(ordering):
    from idmx import configs
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(128, 128), xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0,
                             set_ordering=ordering)
    n0 = configs.droplet2d_noise()
    sim.SetDensityField(n0)
    sim.MainLoop([0, 200])
    sim.ComputeMomentsMeta()
    st = cpuref.State('D2Q9', (128, 128), aos=(ordering == 'cpu'))
    ref0 = cpuref.State('D2Q9', (128, 128))
    cpuref.init_pop(ref0, n0, np.zeros(2 * 128 * 128))
    st.pop[:] = ref0.pop.reshape(9, -1).T.ravel() if ordering == 'cpu' else ref0.pop
    cpuref.meta_run(st, 'D2E4', -5.0 / 3.0, 1.0, nsteps=200)
    cpuref.meta_moments(st, 'D2E4', -5.0 / 3.0)
    assert rel_err(sim.GetDensityField().ravel(), st.n) <= 1e-12
    e_u = rel_err(sim.GetVelocityField().ravel(), st.u)
    print('u rel err', e_u, flush=True)
    assert e_u <= 1e-12
    pops = sim.GetPopulations()
    assert pops.shape == ((128, 128, 9) if ordering == 'cpu' else (9, 128, 128))
    assert rel_err(pops.ravel(), st.pop) <= 1e-12
    sim.End()


def test_meta_mainloop_does_not_write_u_and_checks_run():
    from idmx import configs
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from idmx.LBM.MultiPhaseLoopChecks import CheckUConvergenceSCMP
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(64, 64), xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0)
    sim.SetDensityField(configs.droplet((64, 64), 12.0))
    sim.MainLoop([0, 50])
    assert np.all(sim.sims_idpy_memory['u'].D2H() == 0.0)
    sim.MainLoop([0, 100, 200, 300], convergence_functions=[CheckUConvergenceSCMP])
    assert len(sim.sims_vars['max_u']) == 3
    assert 0 < sim.sims_vars['max_u'][-1] < 0.1
    p_in, p_out, dp = sim.DeltaPLaplace()
    assert dp > 0
    sim.End()


def test_center_of_mass_and_laplace_surface_tension():
    """Device centre of mass equals the numpy one on the same field; the
    Laplace surface tension sigma = dP R_eq is the same for two droplet
    radii once both have relaxed (Laplace's law), and dP > 0."""
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    out = []
    for R in (12.0, 20.0):
        sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(96, 96), xi_stencil=XIStencils['D2Q9'],
                                 f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                                 psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0)
        sim.InitRadialInterface(n_g=0.13, n_l=1.9, R=R)
        sim.MainLoop([0, 20000])
        sim.ComputeMomentsMeta()
        nh = sim.sims_idpy_memory['n'].D2H()
        mass, cm = sim.ComputeCenterOfMass()
        ys, xs = np.divmod(np.arange(96 * 96), 96)
        assert abs(mass - nh.sum()) <= 1e-12 * nh.sum()
        assert abs(cm[0] - (nh * xs).sum() / nh.sum()) <= 1e-10
        assert abs(cm[1] - (nh * ys).sum() / nh.sum()) <= 1e-10
        sigma, R_eq, dP = sim.SurfaceTensionLaplace()
        print('R', R, 'R_eq', R_eq, 'dP', dP, 'sigma', sigma, flush=True)
        assert dP > 0 and abs(R_eq - R) < 2.0
        out.append(sigma)
        sim.End()
    assert abs(out[0] - out[1]) <= 0.05 * out[1], out


def test_dump_pop_snapshot(tmp_path):
    """DumpPopSnapshot of a running simulation holds its device arrays and
    scalar variables under ShanChenMultiPhase/... (IdpySims.py:73-136)."""
    from idmx import configs
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(64, 48), xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0)
    sim.SetDensityField(configs.droplet((64, 48), 10.0))
    sim.MainLoop([0, 50])
    fn = str(tmp_path / 'sim.hdf5')
    sim.DumpPopSnapshot(file_name=fn)
    c = 'ShanChenMultiPhase'
    assert np.array_equal(sim.ReadSnapshotData(fn, c + '/idpy_memory/pop'),
                          sim.sims_idpy_memory['pop'].D2H())
    assert np.array_equal(sim.ReadSnapshotData(fn, c + '/idpy_memory/n'),
                          sim.sims_idpy_memory['n'].D2H())
    assert int(sim.ReadSnapshotData(fn, c + '/vars/V')) == 64 * 48
    sim.End()


@pytest.mark.parametrize('cls', ['legacy', 'meta'])
def test_compute_moments_unshifted_velocity(cls):
    """ComputeMoments (K_ComputeMoments, LBM.py:1667-1698; the current
    class's MultiPhase.py:1543-1563): n = sum_q f_q and u = sum_q xi_q f_q / n
    without the force shift, checked against those sums taken on the host
    over the device populations (SoA, gpu ordering)."""
    from idmx import configs
    from idmx.IdpyCode import HIP_T
    from idmx.LBM import LBM, MultiPhase
    mod = LBM if cls == 'legacy' else MultiPhase
    n = sp.Symbol('n')
    sim = mod.ShanChenMultiPhase(lang=HIP_T, dim_sizes=(48, 40),
                                 xi_stencil=mod.XIStencils['D2Q9'],
                                 f_stencil=mod.FStencils['D2E4'],
                                 psi_code='1. - exp(-ln)' if cls == 'legacy' else '1. - exp(-ln_0)',
                                 psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0)
    sim.SetDensityField(configs.droplet((48, 40), 10.0))
    sim.MainLoop([0, 60])
    sim.ComputeMoments()
    m = sim.sims_idpy_memory
    V = 48 * 40
    f = m['pop'].D2H().reshape(9, V)
    xi = np.array(mod.XIStencils['D2Q9']['XIs'], dtype=np.float64)
    n_h = f.sum(axis=0)
    u_h = (xi.T @ f) / n_h
    n_d, u_d = m['n'].D2H().ravel(), m['u'].D2H().reshape(2, V)
    if cls == 'meta':
        # loops the reference defines no behaviour for raise, not AttributeError
        for loop in ('MainLoopSimpleWallsInletOutlet', 'MainLoopSimpleWallsInletOutletSC',
                     'MainLoopGross2011SRT'):
            with pytest.raises(NotImplementedError):
                getattr(sim, loop)([0, 10])
    sim.End()
    assert rel_err(n_d, n_h) <= 1e-15
    assert np.abs(u_d - u_h).max() <= 1e-15 * max(1.0, np.abs(u_h).max()) + 1e-17
    assert np.abs(u_h).max() > 1e-6


@pytest.mark.parametrize('kind', ['diagonal', 'coupled'])
def test_mainloopmrt_with_relaxation_matrix(kind):
    """MainLoopMRT through the Meta class, with S built by RelaxationMatrix
    as a notebook does: symbolic rates fixed by omega_syms_vals (diagonal:
    moment-space kernel, against the oracle's moment-space evaluation and its
    dense matrices) or a matrix coupling two moments (dense tables, against
    the oracle's dense matrices)."""
    from helpers import droplet_field
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.Collisions import RelaxationMatrix
    from idmx.LBM.Equilibria import HermiteEquilibria
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref, hermite_mrt
    dims = (26, 21, 19)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D3Q19'],
                             f_stencil=FStencils['D3E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0)
    rm = RelaxationMatrix(HermiteEquilibria(xi_stencil=XIStencils['D3Q19']))
    rates = np.random.default_rng(8).uniform(0.7, 1.5, 19)
    om = sp.symbols('omega_0:19')
    S = rm.GetDiagonalMatrix(list(om))
    vals = [(o, sp.Float(v, 17)) for o, v in zip(om, rates)]
    if kind == 'coupled':
        ms = rm.GetMoments()
        S = S + rm.SetOmegasByMoment({(ms[4], ms[7]): sp.Float(0.05, 17)})
    n0 = droplet_field(dims, seed=41)
    sim.SetDensityField(n0)
    sim.MainLoopMRT([0, 20, 60], relaxation_matrix=S, omega_syms_vals=vals)
    f = sim.sims_idpy_memory['pop'].D2H().ravel()
    sim.End()
    Snum = np.array(S.subs(vals).evalf(), dtype=np.float64)
    Ssym = sp.Matrix(19, 19, lambda i, j: sp.Rational(Snum[i, j]).limit_denominator(10 ** 15))
    st = cpuref.State('D3Q19', dims)
    cpuref.init_pop(st, n0, np.zeros(3 * n0.size))
    cpuref.meta_run(st, 'D3E4', -5.0 / 3.0, 1.0, nsteps=60,
                    mrt=hermite_mrt.mrt_matrices('D3Q19', Ssym))
    e_dense = rel_err(f, st.pop)
    e_mom = None
    if kind == 'diagonal':
        st2 = cpuref.State('D3Q19', dims)
        cpuref.init_pop(st2, n0, np.zeros(3 * n0.size))
        cpuref.meta_run_mrt_moment(st2, 'D3E4', -5.0 / 3.0, np.diag(Snum), nsteps=60)
        e_mom = rel_err(f, st2.pop)
        assert e_mom <= 1e-12
    print(kind, 'MainLoopMRT rel err vs dense oracle', e_dense, 'vs moment oracle', e_mom,
          flush=True)
    assert e_dense <= 1e-12


@pytest.mark.parametrize('step_mode', ['fused', 'two_pass'])
def test_mainloop_profiling_per_step_kernel_times(step_mode):
    """MainLoop(profiling=True) returns IdpyLoopProfile's layout: per
    checkpoint, sequence 0 maps each kernel name to one time per step
    (seconds, from HIP events around that step's kernels), plus the device
    name; the fused step is one kernel, the two-pass step two, and the
    per-step times add up to the checkpoint's own event timing."""
    from idmx import configs
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(128, 128), xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0,
                             step_mode=step_mode)
    sim.SetDensityField(configs.droplet2d_noise())
    prof = sim.MainLoop([0, 7, 20], profiling=True)
    total_ms = sim.engine.last_run_ms()
    f_prof = sim.sims_idpy_memory['pop'].D2H().ravel()
    sim.End()
    names = (['K_SCStep'] if step_mode == 'fused'
             else ['K_ComputeDensityPsiMeta', 'K_ForceCollideStreamSCMPMeta'])
    assert sorted(prof) == [7, 20]
    for step, nst in ((7, 7), (20, 13)):
        seq = prof[step][0]
        assert isinstance(seq['device_name'], str)
        assert sorted(k for k in seq if k != 'device_name') == sorted(names)
        for k in names:
            t = seq[k]
            assert t.shape == (nst,) and np.all(t > 0) and np.all(t < 1.0), (k, t)
    per_step = sum(prof[20][0][k].sum() for k in names) * 1e3
    print(step_mode, 'sum of step times', per_step, 'ms, checkpoint', total_ms, 'ms',
          flush=True)
    assert per_step <= total_ms * 1.0001 + 1e-3
    assert per_step >= 0.5 * total_ms
    # the same trajectory as without profiling
    sim2 = ShanChenMultiPhase(lang=HIP_T, dim_sizes=(128, 128), xi_stencil=XIStencils['D2Q9'],
                              f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                              psi_sym=1 - sp.exp(-n), SC_G=-5.0 / 3.0, tau=1.0,
                              step_mode=step_mode)
    sim2.SetDensityField(configs.droplet2d_noise())
    assert sim2.MainLoop([0, 7, 20]) is None
    assert np.array_equal(sim2.sims_idpy_memory['pop'].D2H().ravel(), f_prof)
    sim2.End()
