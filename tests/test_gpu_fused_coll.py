"""The Shan-Chen forcing (MainLoopSimpleSC) and the isotropy filter
(MainLoopIsoFilter) on the fused step against the two-pass step they also
run on (walls, body force, step_mode='two_pass'): the same arithmetic per
site (psi from the same densities, the E-stencil sums in the same order),
so populations and the written n, psi, u agree to rounding."""
import numpy as np
import pytest

from helpers import droplet_field, rel_err

pytestmark = pytest.mark.gpu

CASES = [('D2Q9', 'D2E4', (96, 70), 'sc', 'fp64'),
         ('D2Q9', 'D2E8', (300, 33), 'sc', 'fp64'),
         ('D3Q19', 'D3E4', (40, 20, 18), 'sc', 'fp64'),
         ('D3Q19', 'D3E4', (260, 6, 5), 'sc', 'fp64'),
         ('D2Q9', 'D2E4', (64, 48), 'sc', 'fp32s'),
         ('D2Q9', 'D2E4', (64, 37), 'iso', 'fp64'),
         ('D2Q9', 'D2E8', (260, 40), 'iso', 'fp64')]


def _run(tenet, xi, es, dims, kind, prec, two_pass, nsteps=60):
    from idmx.IdpyCode import IdpyMemory
    from idmx.LBM._base import PopArray
    from idmx.LBM._engine import SCEngine
    G = -5.0 / 3.0
    V = int(np.prod(dims))
    D = len(dims)
    eng = SCEngine(tenet, xi, es, dims, 'exp', G, 1.0, precision=prec, rho0=1.0,
                   forcing='sc' if kind == 'sc' else 'guo', iso_filter=kind == 'iso')
    n = IdpyMemory.OnDevice(droplet_field(dims, min(dims) // 4), tenet)
    psi = IdpyMemory.Zeros(V, np.float64, tenet)
    u = IdpyMemory.Zeros(D * V, np.float64, tenet)
    a, b = PopArray(eng, tenet), PopArray(eng, tenet)
    eng.init_pop(a, n, u)
    if eng.run(a, b, nsteps - 1, n=n, psi=psi, u=u, two_pass=two_pass):
        a, b = b, a
    if eng.run(a, b, 1, n=n, psi=psi, u=u, write_fields=True, two_pass=two_pass):
        a, b = b, a
    ms = eng.last_run_ms()
    out = (eng.pop_to_host_soa(a).ravel(), n.D2H().copy(), psi.D2H().copy(), u.D2H().copy())
    names = eng.step_kernel_names(two_pass)
    for x in (n, psi, u, a, b):
        x.Free()
    eng.close()
    return out, names


@pytest.mark.parametrize('xi,es,dims,kind,prec', CASES)
def test_fused_step_equals_two_pass(tenet, xi, es, dims, kind, prec):
    (fa, na, pa, ua), names_f = _run(tenet, xi, es, dims, kind, prec, False)
    (fb, nb, pb, ub), names_t = _run(tenet, xi, es, dims, kind, prec, True)
    tol = 1e-14 if prec == 'fp64' else 1e-6
    errs = [rel_err(fa, fb), rel_err(na, nb), rel_err(pa, pb), rel_err(ua, ub)]
    print(xi, es, dims, kind, prec, 'pop/n/psi/u rel err', ['%.1e' % e for e in errs],
          'kernels', names_f, names_t, flush=True)
    assert names_f == (None, 'K_SCStep')
    assert names_t[1] == 'K_ForceCollideStreamSCMPMeta'
    assert max(errs) <= tol
