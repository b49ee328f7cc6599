"""HIP path vs the CPU oracle on the same seeded inputs (FP64, 1e-12).

The HIP kernels are reached through the C ABI (SCEngine wraps idmx_sc_*);
the oracle is oracle/libcpuref.so. Sizes are small and deliberately not
multiples of the tile shapes, so partial tiles and periodic wrap of the halo
are exercised.
"""
import zlib

import numpy as np
import pytest

from helpers import droplet_field, rel_err

G = -5.0 / 3.0

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = [
    ('D2Q9', 'D2E4', (37, 29)),
    ('D2Q9', 'D2E6', (40, 23)),
    ('D2Q9', 'D2E8', (300, 17)),
    ('D2Q9', 'D2E8', (13, 70)),
    ('D3Q19', 'D3E4', (13, 11, 9)),
    ('D3Q19', 'D3E6', (35, 10, 7)),
    ('D3Q19', 'D3E8', (9, 12, 11)),
    ('D3Q19', 'D3E4', (70, 9, 5)),
]


def _device_fields(tenet, eng, n0):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    V, D = eng.V, eng.D
    n = IdpyMemory.OnDevice(n0, tenet)
    u = IdpyMemory.Zeros(V * D, np.float64, tenet)
    psi = IdpyMemory.Zeros(V, np.float64, tenet)
    pop = PopArray(eng, tenet)
    swap = PopArray(eng, tenet)
    eng.init_pop(pop, n, u)
    return pop, swap, n, psi, u


def _oracle(xi, es, dims, n0, nsteps, G=G, tau=1.0, mrt=None, rates=None):
    """The oracle's Meta step: BGK, MRT with dense matrices (mrt), or MRT
    with a diagonal S evaluated in moment space (rates)."""
    from oracle import cpuref
    V, D = int(np.prod(dims)), len(dims)
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0, np.zeros(V * D))
    if rates is not None:
        cpuref.meta_run_mrt_moment(st, es, G, rates, nsteps=nsteps)
    else:
        cpuref.meta_run(st, es, G, tau, nsteps=nsteps, mrt=mrt)
    n1, u1, psi1 = st.n.copy(), None, st.psi.copy()
    mom = cpuref.State(xi, dims)
    mom.pop[:] = st.pop
    cpuref.meta_moments(mom, es, G)
    return st.pop.copy(), n1, psi1, mom


@pytest.mark.parametrize('xi,es,dims', CASES)
@pytest.mark.parametrize('mode', ['fused', 'two_pass'])
def test_bgk_fp64_matches_oracle(tenet, xi, es, dims, mode):
    from idmx.LBM._engine import SCEngine
    print('case', xi, es, dims, mode, flush=True)
    n0 = droplet_field(dims, seed=zlib.crc32(repr((xi, es, dims)).encode()) % 1000)
    nsteps = 40
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0)
    pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
    if eng.run(pop, swap, nsteps, n=n, psi=psi, u=u, write_fields=True,
               two_pass=(mode == 'two_pass')):
        pop, swap = swap, pop
    f_gpu = eng.pop_to_host_soa(pop).ravel()
    f_ref, n_ref, psi_ref, mom = _oracle(xi, es, dims, n0, nsteps)
    assert rel_err(f_gpu, f_ref) <= TOL
    # n / psi of the start of the last step, as the reference's arrays hold
    assert rel_err(n.D2H(), n_ref) <= TOL
    assert rel_err(psi.D2H(), psi_ref) <= TOL
    # ComputeMomentsMeta on the final state
    from idmx.IdpyCode import IdpyMemory
    n2 = IdpyMemory.Zeros(eng.V, np.float64, tenet)
    p2 = IdpyMemory.Zeros(eng.V, np.float64, tenet)
    u2 = IdpyMemory.Zeros(eng.V * eng.D, np.float64, tenet)
    eng.moments(pop, n2, p2, u2)
    assert rel_err(n2.D2H(), mom.n) <= TOL
    assert rel_err(p2.D2H(), mom.psi) <= TOL
    e_u = rel_err(u2.D2H(), mom.u)
    print('u rel err', xi, es, dims, mode, e_u, flush=True)
    assert e_u <= 1e-12
    eng.close()


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (33, 31)),
                                        ('D3Q19', 'D3E4', (12, 13, 10))])
def test_mass_conserved_to_roundoff(tenet, xi, es, dims):
    from idmx.LBM._engine import SCEngine
    n0 = droplet_field(dims, seed=7)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0)
    pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
    m0, j0 = eng.mass_momentum(pop)
    if eng.run(pop, swap, 500):
        pop, swap = swap, pop
    m1, j1 = eng.mass_momentum(pop)
    assert abs(m1 - m0) / m0 < 5e-14
    # the Shan-Chen force sums to zero on a periodic lattice: momentum too
    assert np.abs(j1 - j0).max() < 1e-11 * m0
    eng.close()


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (24, 21)),
                                        ('D3Q19', 'D3E4', (11, 10, 9))])
def test_mrt_matches_oracle(tenet, xi, es, dims):
    """MRT with a diagonal S of distinct rates: HIP tables from the package's
    rational Hermite basis, oracle matrices from its own sympy derivation."""
    from idmx.LBM import hermite
    from idmx.LBM._engine import SCEngine
    from oracle import hermite_mrt
    Q = 9 if xi == 'D2Q9' else 19
    rng = np.random.default_rng(3)
    s = rng.uniform(0.6, 1.6, size=Q)
    S = np.diag(s)
    n0 = droplet_field(dims, seed=11)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0, mrt=hermite.mrt_tables(xi, S))
    pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
    if eng.run(pop, swap, 30):
        pop, swap = swap, pop
    import sympy as sp
    Ssym = sp.diag(*[sp.Rational(v).limit_denominator(10 ** 15) for v in s])
    mats = hermite_mrt.mrt_matrices(xi, Ssym)
    f_ref, _, _, _ = _oracle(xi, es, dims, n0, 30, mrt=mats)
    err = rel_err(eng.pop_to_host_soa(pop).ravel(), f_ref)
    print('dense MRT rel err', err, flush=True)
    assert err <= TOL
    eng.close()


def test_fp32_shifted_close_to_fp64(tenet):
    from idmx.LBM._engine import SCEngine
    dims = (64, 48)
    n0 = droplet_field(dims, seed=5)
    out = {}
    for prec in ('fp64', 'fp32s'):
        eng = SCEngine(tenet, 'D2Q9', 'D2E4', dims, 'exp', G, 1.0, precision=prec,
                       rho0=1.0)
        pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
        if eng.run(pop, swap, 100):
            pop, swap = swap, pop
        out[prec] = eng.pop_to_host_soa(pop)
        eng.close()
    assert rel_err(out['fp32s'], out['fp64']) < 2e-5


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (37, 29)), ('D2Q9', 'D2E8', (300, 17)),
                                        ('D3Q19', 'D3E4', (70, 19, 9)),
                                        ('D3Q19', 'D3E8', (35, 17, 7))])
def test_every_launch_variant_matches_oracle(tenet, xi, es, dims):
    """Each tile shape, march length and register/re-read variant the
    autotuner may pick gives the oracle's result."""
    from idmx import tune
    from idmx.LBM._engine import SCEngine
    n0 = droplet_field(dims, seed=21)
    f_ref, _, _, _ = _oracle(xi, es, dims, n0, 12)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0)
    cands = [c for c in tune.sim_candidates(len(dims), dims[-1]) if eng.launch_ok(*c)]
    # the tuner's march lengths exceed these small domains, so also march
    # short segments (segment boundaries, a last shorter segment)
    for cand in cands + [c[:2] + (3,) + c[3:] for c in cands] + [
            c[:2] + (4,) + c[3:] for c in cands[::5]]:
        print('variant', cand, flush=True)
        eng.set_launch(*cand)
        pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
        if eng.run(pop, swap, 12):
            pop, swap = swap, pop
        assert rel_err(eng.pop_to_host_soa(pop).ravel(), f_ref) <= TOL, cand
        for x in (pop, swap, n, psi, u):
            x.Free()
    eng.close()


@pytest.mark.parametrize('xi,es,dims,prec,coll', [
    ('D3Q19', 'D3E4', (70, 19, 9), 'fp64', 'diag'), ('D3Q19', 'D3E4', (70, 19, 9), 'fp64', 'dense'),
    ('D3Q19', 'D3E4', (70, 19, 9), 'fp32s', 'diag'), ('D3Q19', 'D3E4', (71, 19, 9), 'fp32s', 'bgk'),
    ('D3Q19', 'D3E8', (35, 17, 7), 'fp64', 'diag'), ('D2Q9', 'D2E8', (300, 17), 'fp64', 'diag'),
    ('D2Q9', 'D2E4', (37, 29), 'fp32s', 'bgk')])
def test_lds_staged_variant_matches_oracle(tenet, xi, es, dims, prec, coll):
    """Variants 9 and 10 (lead plane copied into LDS by DMA while the
    previous plane collides; 10 on the x-shifted storage) with BGK, diagonal
    and dense MRT, FP64 and FP32-shifted, on sizes that leave partial tiles
    and a short last segment, against the oracle; every tile shape whose
    staging buffer fits a CU."""
    import sympy as sp
    from idmx import tune
    from idmx.LBM import hermite
    from idmx.LBM._engine import SCEngine
    from oracle import hermite_mrt
    Q = 9 if xi == 'D2Q9' else 19
    mats, tables = None, None
    if coll != 'bgk':
        rng = np.random.default_rng(6)
        s = rng.uniform(0.6, 1.6, size=Q)
        # dense: the (A, beta, gamma) tables of the same diagonal S
        tables = (hermite.mrt_kernel_tables(xi, np.diag(s)) if coll == 'diag'
                  else hermite.mrt_tables(xi, np.diag(s)))
        assert isinstance(tables, dict) == (coll == 'diag')
        mats = hermite_mrt.mrt_matrices(
            xi, sp.diag(*[sp.Rational(v).limit_denominator(10 ** 15) for v in s]))
    n0 = droplet_field(dims, seed=33)
    if coll == 'diag':
        f_ref, _, _, _ = _oracle(xi, es, dims, n0, 12, rates=s)
    else:
        f_ref, _, _, _ = _oracle(xi, es, dims, n0, 12, mrt=mats)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0, mrt=tables, precision=prec, rho0=1.0)
    cands = [c for c in tune.sim_candidates(len(dims), dims[-1]) if c[3] in (9, 10)]
    cands += [c[:2] + (3, 9) for c in cands]
    tol = TOL if prec == 'fp64' else 5e-6
    ran = 0
    for cand in cands:
        if not eng.launch_ok(*cand):
            continue
        print('variant', cand, flush=True)
        eng.set_launch(*cand)
        pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
        if eng.run(pop, swap, 12):
            pop, swap = swap, pop
        assert rel_err(eng.pop_to_host_soa(pop).ravel(), f_ref) <= tol, cand
        for x in (pop, swap, n, psi, u):
            x.Free()
        ran += 1
    eng.close()
    assert ran >= 4


@pytest.mark.parametrize('xi,es,dims,prec', [('D2Q9', 'D2E4', (24, 21), 'fp64'),
                                             ('D3Q19', 'D3E4', (11, 10, 9), 'fp64'),
                                             ('D3Q19', 'D3E8', (33, 17, 9), 'fp64'),
                                             ('D3Q19', 'D3E4', (11, 10, 9), 'fp32s')])
def test_mrt_diagonal_path_matches_oracle(tenet, xi, es, dims, prec):
    """The moment-space MRT kernel (diagonal S, compiled-in bases) against the
    oracle's dense matrices from its own sympy derivation, and against the
    oracle's moment-space evaluation of the same sums."""
    import sympy as sp
    from idmx.LBM import hermite
    from idmx.LBM._engine import SCEngine
    from oracle import hermite_mrt
    Q = 9 if xi == 'D2Q9' else 19
    rng = np.random.default_rng(4)
    s = rng.uniform(0.6, 1.6, size=Q)
    tables = hermite.mrt_kernel_tables(xi, np.diag(s))
    assert 'diag' in tables
    n0 = droplet_field(dims, seed=12)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0, mrt=tables, precision=prec, rho0=1.0)
    pop, swap, n, psi, u = _device_fields(tenet, eng, n0)
    if eng.run(pop, swap, 30):
        pop, swap = swap, pop
    Ssym = sp.diag(*[sp.Rational(v).limit_denominator(10 ** 15) for v in s])
    mats = hermite_mrt.mrt_matrices(xi, Ssym)
    f_ref, _, _, _ = _oracle(xi, es, dims, n0, 30, mrt=mats)
    f_mom, _, _, _ = _oracle(xi, es, dims, n0, 30, rates=s)
    f_gpu = eng.pop_to_host_soa(pop).ravel()
    e_dense, e_mom = rel_err(f_gpu, f_ref), rel_err(f_gpu, f_mom)
    print('diagonal MRT rel err vs dense oracle', e_dense, 'vs moment oracle', e_mom, flush=True)
    tol = TOL if prec == 'fp64' else 5e-6
    assert e_dense <= tol and e_mom <= tol
    eng.close()
