"""idmx_batch_force_rest: a domain's populations at rest under the Guo
forcing. The oracle's moments (its own Shan-Chen force) see density n and
velocity (sum xi f + F/2) / n = 0 on every site; a domain set back to its
own converged density moves far less in its first steps with these
populations than with f = w n, which starts every interface at F/(2n)
(what remains is the steady state's small spurious current)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('prec', ['fp64', 'fp32s'])
@pytest.mark.parametrize('es', ['E4', 'E6', 'E8'])
def test_force_rest_moments_match_oracle(tenet, es, prec):
    from idmx import configs, sweep
    from oracle import cpuref
    pts = [sweep.SweepPoint(L=48, G=-5.0 / 3.0, R=9.0), sweep.SweepPoint(L=33, G=-1.5, R=7.5)]
    b = sweep.Batch(tenet, pts, prec, es)
    rng = np.random.default_rng(7)
    fields = []
    for i, p in enumerate(pts):
        # off-centre droplet plus noise: no symmetry that would hide a
        # transposed axis or a sign
        n = configs.droplet((p.L, p.L), p.R, 2.0, centre=(0.4 * p.L, 0.55 * p.L))
        n = n + rng.uniform(-1e-3, 1e-3, n.size)
        b.set_density(i, n)
        fields.append(n)
    b.force_rest(-1)
    tol = 1e-13 if prec == 'fp64' else 2e-6
    for i, p in enumerate(pts):
        f = b.get_populations(i)
        st = cpuref.State('D2Q9', (p.L, p.L))
        st.pop[:] = f.ravel()
        cpuref.meta_moments(st, 'D2' + es, p.G)
        e_n = float(np.abs(st.n - fields[i]).max() / np.abs(fields[i]).max())
        u_max = float(np.abs(st.u).max())
        print(es, prec, i, 'n rel err %.2e max |u| %.2e' % (e_n, u_max), flush=True)
        assert e_n <= tol
        assert u_max <= tol
    b.close()


def test_force_rest_restarts_at_steady_state(tenet):
    """A 64^2 droplet converged to max |dn| < 1e-12, its density written
    back: the first 200 steps move n at least five times less after
    force_rest than from f = w n."""
    from idmx import sweep
    p = sweep.SweepPoint(L=64, G=-1.7, R=16.0)
    b = sweep.Batch(tenet, [p], 'fp64', 'E4')
    b.set_density(0, sweep.cold_start(p))
    b.check()
    for _ in range(2000):
        b.run(200)
        if b.check()[0] < 1e-12:
            break
    n_eq = b.get_density(0).ravel()
    kick = []
    for rest in (False, True):
        b.set_density(0, n_eq)
        if rest:
            b.force_rest(0)
        b.run(200)
        kick.append(float(b.check()[0]))
    b.close()
    print('max |dn| after 200 steps: f = w n %.2e, force_rest %.2e' % tuple(kick), flush=True)
    assert kick[1] < 0.2 * kick[0]
