"""Edge shapes of the lattice: every axis at the smallest size the C ABI
accepts (2 R + 1, so the halo of a tile wraps onto the tile itself), and
ragged boxes with one axis at that size and another long, against the
oracle (FP64, 1e-12): the default fused launch (writing n / psi on the last
step), the two-pass step, and the LDS-staged step (3D 64 x 4 tiles on the
streamed and on the x-shifted storage, 2D rows of 64 on the x-shifted
storage)."""
import numpy as np
import pytest

from helpers import rel_err, smooth_field

pytestmark = pytest.mark.gpu

G = -1.0

EDGES = [
    ('D2Q9', 'D2E4', (3, 3)), ('D2Q9', 'D2E8', (5, 5)), ('D2Q9', 'D2E6', (5, 5)),
    ('D2Q9', 'D2E4', (3, 300)), ('D2Q9', 'D2E4', (300, 3)), ('D2Q9', 'D2E8', (5, 70)),
    ('D3Q19', 'D3E4', (3, 3, 3)), ('D3Q19', 'D3E8', (5, 5, 5)), ('D3Q19', 'D3E6', (5, 5, 5)),
    ('D3Q19', 'D3E4', (3, 40, 3)), ('D3Q19', 'D3E4', (70, 3, 5)),
]


@pytest.mark.parametrize('xi,es,dims', EDGES)
def test_smallest_and_ragged_lattices_match_oracle(tenet, xi, es, dims):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    from oracle import cpuref
    print('case', xi, es, dims, flush=True)
    n0 = smooth_field(dims, seed=sum(dims))
    nsteps = 20
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0, np.zeros(st.V * st.D))
    m_ref0 = st.pop.sum()
    cpuref.meta_run(st, es, G, 1.0, nsteps=nsteps)
    modes = [('fused', None), ('two_pass', None)]
    if xi == 'D3Q19':
        modes += [('fused', (64, 4, 8, 9)), ('fused', (64, 4, 8, 10))]
    else:
        modes.append(('fused', (64, 1, 8, 10)))
    for mode, launch in modes:
        eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0)
        if launch is not None:
            assert eng.launch_ok(*launch)
            eng.set_launch(*launch)
        n = IdpyMemory.OnDevice(n0, tenet)
        u = IdpyMemory.Zeros(eng.V * eng.D, np.float64, tenet)
        psi = IdpyMemory.Zeros(eng.V, np.float64, tenet)
        pop, swap = PopArray(eng, tenet), PopArray(eng, tenet)
        eng.init_pop(pop, n, u)
        if eng.run(pop, swap, nsteps, n=n, psi=psi, u=u, write_fields=True,
                   two_pass=(mode == 'two_pass')):
            pop, swap = swap, pop
        f = eng.pop_to_host_soa(pop).ravel()
        e_pop = rel_err(f, st.pop)
        e_n = rel_err(n.D2H(), st.n)
        e_psi = rel_err(psi.D2H(), st.psi)
        for x in (n, u, psi, pop, swap):
            x.Free()
        eng.close()
        print(mode, launch, 'rel err pop', e_pop, 'n', e_n, 'psi', e_psi, flush=True)
        assert e_pop <= 1e-12 and e_n <= 1e-12 and e_psi <= 1e-12
        assert abs(f.sum() - m_ref0) <= 1e-13 * m_ref0
