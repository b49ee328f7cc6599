"""The fused step with walls (the lead plane, the caller's psi plane and the
wall masks staged in LDS by DMA; fluid sites only, half-way bounce-back,
optional body force) against the two-pass walls step on the same input:
the same arithmetic in the same order, so the fluid sites agree to
rounding. Solid blocks cross the periodic boundary and the tiles are
partial. The fused steps of a run keep the populations in the x-shifted
storage between them; a run that writes the fields ends with the two-pass
step, one that does not with a fused step that writes the streamed storage
(both are covered)."""
import numpy as np
import pytest

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

G = -5.0 / 3.0


def _walls(dims, seed):
    rng = np.random.default_rng(seed)
    w = np.ones(tuple(reversed(dims)), dtype=np.uint8)
    if len(dims) == 2:
        w[0, :] = 0
        w[-1, :] = 0
        w[dims[1] // 3:dims[1] // 3 + 3, :3] = 0   # a block across x = 0
        w[dims[1] // 3:dims[1] // 3 + 3, -2:] = 0
        w[dims[1] // 2:dims[1] // 2 + 4, dims[0] // 2:dims[0] // 2 + 5] = 0
    else:
        w[:, 0, :] = 0
        w[:, -1, :] = 0
        w[2:5, 4:8, :2] = 0                        # across x = 0 and z = 0
        w[-2:, 4:8, -3:] = 0
        w[dims[2] // 2:dims[2] // 2 + 3, 6:10, 10:16] = 0
    psi_w = np.where(w == 0, rng.uniform(0.3, 0.7, size=w.shape), 0.0)
    return w.ravel(), psi_w.ravel()


@pytest.mark.parametrize('xi,es,dims,g,nsteps', [('D2Q9', 'D2E4', (300, 37), 0.0, 60),
                                                 ('D2Q9', 'D2E4', (70, 33), 2e-5, 60),
                                                 ('D3Q19', 'D3E4', (70, 19, 13), 0.0, 30),
                                                 ('D3Q19', 'D3E4', (35, 17, 9), 1e-5, 30)])
@pytest.mark.parametrize('wf', [True, False])
def test_fused_walls_step_equals_two_pass(tenet, xi, es, dims, g, nsteps, wf):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    w, psi_w = _walls(dims, 3)
    n0 = droplet_field(dims, seed=8)
    D = len(dims)
    V = int(np.prod(dims))
    wd = IdpyMemory.OnDevice(w, tenet)
    out = {}
    for mode in ('fused', 'two_pass'):
        eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0)
        eng.set_walls(wd)
        if g:
            eng.set_body_force(0.0, g, 0.0)
        n = IdpyMemory.OnDevice(n0, tenet)
        psi = IdpyMemory.OnDevice(psi_w, tenet)
        u = IdpyMemory.Zeros(D * V, np.float64, tenet)
        a, b = PopArray(eng, tenet), PopArray(eng, tenet)
        eng.init_pop(a, n, u)
        if eng.run(a, b, nsteps, n=n, psi=psi, u=u, write_fields=wf,
                   two_pass=mode == 'two_pass', fused_walls=mode == 'fused'):
            a, b = b, a
        ms = eng.last_run_ms()
        out[mode] = (eng.pop_to_host_soa(a).reshape(-1, V), n.D2H(), psi.D2H(), u.D2H(), ms)
        for x in (n, psi, u, a, b):
            x.Free()
        eng.close()
    wd.Free()
    fl = w == 1
    f1, n1, p1, u1, t1 = out['fused']
    f2, n2, p2, u2, t2 = out['two_pass']
    e_pop = rel_err(f1[:, fl].ravel(), f2[:, fl].ravel())
    print(xi, es, dims, 'g', g, 'pop', e_pop, 'identical', np.array_equal(f1[:, fl], f2[:, fl]),
          'ms per step fused', t1 / nsteps, 'two-pass', t2 / nsteps, flush=True)
    assert e_pop <= 1e-15
    if not wf:
        return
    assert rel_err(n1[fl], n2[fl]) <= 1e-15 and rel_err(p1[fl], p2[fl]) <= 1e-15
    assert rel_err(u1.reshape(D, V)[:, fl].ravel(), u2.reshape(D, V)[:, fl].ravel()) <= 1e-14
    assert np.array_equal(p1[~fl], psi_w[~fl])


def test_mainloopwalls_large_lattice_uses_fused_step(tenet):
    """MainLoopWalls through the class on a lattice of more than 2^19 sites,
    where the fused step with walls runs by default, against the oracle
    (fluid sites, 1e-12), and its per-step profile names the fused kernel."""
    import sympy as sp
    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import FStencils, ShanChenMultiPhase, XIStencils
    from oracle import cpuref
    dims = (1024, 520)
    w, psi_w = _walls(dims, 4)
    n0 = droplet_field(dims, seed=12)
    n = sp.Symbol('n')
    sim = ShanChenMultiPhase(lang=HIP_T, dim_sizes=dims, xi_stencil=XIStencils['D2Q9'],
                             f_stencil=FStencils['D2E4'], psi_code='1. - exp(-ln_0)',
                             psi_sym=1 - sp.exp(-n), SC_G=G, tau=1.0)
    sim.SetDensityField(n0)
    sim.GetWalls(w.reshape(dims[::-1]), psi_w.reshape(dims[::-1]))
    prof = sim.MainLoopWalls([0, 20, 40], profiling=True)
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0, np.zeros(2 * st.V))
    st.psi[:] = psi_w
    cpuref.meta_run(st, 'D2E4', G, 1.0, nsteps=40, walls=w)
    fl = w == 1
    pop = sim.sims_idpy_memory['pop'].D2H().reshape(9, -1)
    e_pop = rel_err(pop[:, fl].ravel(), st.pop.reshape(9, -1)[:, fl].ravel())
    e_psi = rel_err(sim.sims_idpy_memory['psi'].D2H()[fl], st.psi[fl])
    seq = prof[40][0]
    sim.End()
    print('large MainLoopWalls pop', e_pop, 'psi', e_psi, 'step ms',
          seq['K_SCStep'][:3] * 1e3, flush=True)
    assert e_pop <= 1e-12 and e_psi <= 1e-12
    assert sorted(k for k in seq if k != 'device_name') == ['K_ComputeDensityPsiMeta', 'K_SCStep']
    # fused steps: no separate density kernel except in the last, two-pass
    # step; the first interval of a fused step is only the gap between two
    # event records (~5 us, up to ~8 us on a single step here), so the
    # medians are compared, not each step against a ~35 us step
    assert np.median(seq['K_ComputeDensityPsiMeta'][:-1]) < \
        0.2 * np.median(seq['K_SCStep'][:-1])
