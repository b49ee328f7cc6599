"""Variant 10 of the fused step (the LDS-staged step on x-shifted storage:
population q of site x kept at x - c_qx while the steps run, so that every
push store covers the tile's own, line-aligned x range) against variant 9
(the same kernel on the streamed storage): the same arithmetic per site, so
populations and the written n, psi, u are bit-identical. Covers odd and even
step counts, the step that writes the fields (it runs on the streamed
storage after the conversion back), partial and wrapped tiles, E8, FP32
shifted storage and D2Q9; the oracle comparison of the variant is in
test_gpu_parity.py."""
import numpy as np
import pytest

from helpers import droplet_field

pytestmark = pytest.mark.gpu

CASES = [('D3Q19', 'D3E4', (70, 19, 9), (64, 4, 8), 'fp64', 'mrt'),
         ('D3Q19', 'D3E4', (13, 11, 9), (64, 4, 8), 'fp64', 'mrt'),
         ('D3Q19', 'D3E4', (96, 64, 40), (64, 4, 16), 'fp64', 'bgk'),
         ('D3Q19', 'D3E4', (71, 19, 9), (64, 4, 8), 'fp32s', 'mrt'),
         ('D3Q19', 'D3E4', (71, 19, 9), (128, 2, 8), 'fp32s', 'mrt'),
         ('D3Q19', 'D3E8', (33, 17, 9), (32, 8, 8), 'fp64', 'bgk'),
         ('D2Q9', 'D2E8', (300, 17), (256, 1, 8), 'fp64', 'bgk'),
         ('D2Q9', 'D2E4', (37, 29), (64, 1, 8), 'fp64', 'bgk')]


def _run(tenet, xi, es, dims, launch, steps, prec, coll, wf):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM import hermite
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    S = np.diag(np.linspace(0.7, 1.5, 19 if xi == 'D3Q19' else 9))
    mrt = hermite.mrt_kernel_tables(xi, S) if coll == 'mrt' else None
    eng = SCEngine(tenet, xi, es, dims, 'exp', -5.0 / 3.0, 1.0, mrt=mrt, precision=prec,
                   rho0=1.0 if prec == 'fp32s' else 0.0)
    if not eng.launch_ok(*launch):
        eng.close()
        return None
    eng.set_launch(*launch)
    assert eng.get_launch() == launch
    V, D = eng.V, eng.D
    n = IdpyMemory.OnDevice(droplet_field(dims, seed=33), tenet)
    u = IdpyMemory.Zeros(V * D, np.float64, tenet)
    psi = IdpyMemory.Zeros(V, np.float64, tenet)
    pop = PopArray(eng, tenet)
    swap = PopArray(eng, tenet)
    eng.init_pop(pop, n, u)
    outs = []
    for st in steps:
        kw = dict(n=n, psi=psi, u=u, write_fields=True) if wf else {}
        if eng.run(pop, swap, st, **kw):
            pop, swap = swap, pop
        outs.append(eng.pop_to_host_soa(pop).copy())
        if wf:
            outs += [n.D2H().copy(), psi.D2H().copy(), u.D2H().copy()]
    for x in (pop, swap, n, u, psi):
        x.Free()
    eng.close()
    return outs


@pytest.mark.parametrize('xi,es,dims,shape,prec,coll', CASES)
@pytest.mark.parametrize('wf,steps', [(False, (1, 2, 7)), (True, (1, 4))])
def test_xshift_matches_variant9(tenet, xi, es, dims, shape, prec, coll, wf, steps):
    print('case', xi, es, dims, shape, prec, coll, wf, steps, flush=True)
    a = _run(tenet, xi, es, dims, shape + (9,), steps, prec, coll, wf)
    b = _run(tenet, xi, es, dims, shape + (10,), steps, prec, coll, wf)
    assert a is not None and b is not None, 'launch shape does not fit'
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), (k, float(np.max(np.abs(x - y))))
