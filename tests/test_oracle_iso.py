"""Pins of the oracle's isotropy filter and of the MainLoopIsoFilter step
(K_IsotropyFilter, idpy/LBM/LBMKernelsMeta.py:895-1080; MultiPhase.py:
1099-1223; no GPU). The filter is checked against the reference's own
construction (a symbolic solve with the moment substitutions), against a
dense numpy solve per site, against the moments it must keep and set
(computed from raw lattice sums), and through properties the construction
fixes: degree-one homogeneity, the rest state as a fixed point, the
Maxwellian third moments of a filtered equilibrium, and F(F(F(f))) =
F(F(f)) although F(F(f)) != F(f)."""
import numpy as np
import pytest

from helpers import droplet_field, rel_err

G = -5.0 / 3.0
XI = np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)],
              dtype=np.float64)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _pops(n, seed, amp=0.2):
    rng = np.random.default_rng(seed)
    return W * rng.uniform(0.3, 2.0, size=(n, 1)) * (1 + amp * rng.standard_normal((n, 9)))


def _raw(f, px, py):
    return f @ (XI[:, 0] ** px * XI[:, 1] ** py)


def test_package_and_oracle_tables_agree_and_invert():
    """Two independent exact derivations (sympy in the oracle, fractions in
    the package) give the same doubles, and Ai inverts the system."""
    import sympy as sp
    from idmx.LBM import hermite
    from oracle import iso_filter
    Mp, Aip = hermite.iso_filter_tables()
    Mo, Aio = iso_filter.tables()
    assert np.array_equal(Mp, Mo) and np.array_equal(Aip, Aio)
    _, A, Ai = iso_filter.exact_tables()
    assert A * Ai == sp.eye(9)


def test_filter_matches_the_reference_symbolic_solve():
    from oracle import cpuref, iso_filter
    f = _pops(64, seed=3)
    g = cpuref.iso_filter(f)
    fn, order = iso_filter.symbolic_solution()
    assert order == ['f_%d' % q for q in range(9)]
    M, _ = iso_filter.tables()
    m = f @ M.T
    m[:, 1:] /= m[:, :1]
    ref = np.array([np.broadcast_to(v, (64,)) for v in fn(*[m[:, k] for k in range(1, 9)])]).T
    assert rel_err(g, ref * m[:, :1]) < 2e-15


def test_filter_matches_a_dense_numpy_solve():
    from oracle import cpuref, np_ref
    f = _pops(200, seed=4)
    assert rel_err(cpuref.iso_filter(f), np_ref.iso_filter_solve(f.T.copy()).T) < 1e-14


def test_filter_keeps_and_sets_the_moments():
    from oracle import cpuref
    f = _pops(100, seed=5)
    g = cpuref.iso_filter(f)
    n = _raw(f, 0, 0)
    ux, uy = _raw(f, 1, 0) / n, _raw(f, 0, 1) / n
    pxx, pyy, pxy = _raw(f, 2, 0) / n, _raw(f, 0, 2) / n, _raw(f, 1, 1) / n
    close = lambda a, b: np.abs(a - b).max() <= 1e-14 * max(1.0, np.abs(b).max())
    # kept: density, momentum, xx and the fourth-order xxyy Hermite moment
    assert close(_raw(g, 0, 0), n)
    assert close(_raw(g, 1, 0), n * ux) and close(_raw(g, 0, 1), n * uy)
    assert close(_raw(g, 2, 0), n * pxx)
    h4 = lambda h: _raw(h, 2, 2) - (_raw(h, 2, 0) + _raw(h, 0, 2)) / 3 + _raw(h, 0, 0) / 9
    assert close(h4(g), h4(f))
    # set from the others (LBMKernelsMeta.py:1003-1015)
    assert close(_raw(g, 1, 1), n * ux * uy)
    assert close(_raw(g, 2, 0) - _raw(g, 0, 2), n * (ux ** 2 - uy ** 2))
    assert close(_raw(g, 2, 1), n * (2 * pxy * ux + pxx * uy - 2 * uy * ux ** 2))
    assert close(_raw(g, 1, 2), n * (2 * pxy * uy + pyy * ux - 2 * ux * uy ** 2))


def test_filter_is_homogeneous_and_fixes_the_rest_state():
    from oracle import cpuref
    f = _pops(50, seed=6)
    assert rel_err(cpuref.iso_filter(3.0 * f), 3.0 * cpuref.iso_filter(f)) < 1e-15
    rest = W * np.linspace(0.1, 2.5, 7)[:, None]
    assert rel_err(cpuref.iso_filter(rest), rest) < 1e-15


def test_filtered_equilibrium_has_maxwellian_third_moments():
    """A second-order D2Q9 equilibrium lacks the n u_x^2 u_y part of the
    continuum third moment (its sum xi_x^2 xi_y f is n u_y / 3); the filter
    restores n (u_y / 3 + u_x^2 u_y) and keeps the equilibrium's n, u and
    second moments."""
    from oracle import cpuref
    rng = np.random.default_rng(7)
    n = rng.uniform(0.2, 2.0, size=(40, 1))
    u = rng.uniform(-0.1, 0.1, size=(40, 2))
    xu = u @ XI.T
    uu = (u ** 2).sum(1, keepdims=True)
    feq = W * n * (1 + 3 * xu + 4.5 * xu ** 2 - 1.5 * uu)
    n = n[:, 0]
    assert np.abs(_raw(feq, 2, 1) - n * u[:, 1] / 3).max() < 1e-15
    g = cpuref.iso_filter(feq)
    assert np.abs(_raw(g, 2, 1) - n * (u[:, 1] / 3 + u[:, 0] ** 2 * u[:, 1])).max() < 1e-14
    assert np.abs(_raw(g, 1, 2) - n * (u[:, 0] / 3 + u[:, 1] ** 2 * u[:, 0])).max() < 1e-14
    for px, py in ((0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 1)):
        assert np.abs(_raw(g, px, py) - _raw(feq, px, py)).max() < 1e-14


def test_filter_squared_is_idempotent():
    from oracle import cpuref
    f = _pops(80, seed=8)
    f1 = cpuref.iso_filter(f)
    f2 = cpuref.iso_filter(f1)
    f3 = cpuref.iso_filter(f2)
    assert rel_err(f2, f1) > 1e-6
    assert rel_err(f3, f2) < 1e-14


@pytest.mark.parametrize('es,dims', [('D2E4', (16, 12)), ('D2E8', (13, 11))])
def test_iso_loop_matches_numpy_brute_force(es, dims):
    from oracle import cpuref, np_ref
    n0 = droplet_field(dims, seed=2)
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0, np.zeros(2 * n0.size))
    f = st.pop.reshape((9,) + tuple(np.flip(dims))).copy()
    for _ in range(5):
        f, n, psi, u = np_ref.step(f, 'D2Q9', es, G, 1.0, iso=True)
    cpuref.iso_run(st, es, G, 1.0, nsteps=5)
    assert rel_err(st.pop, f.ravel()) < 1e-13
    assert rel_err(st.n, n.ravel()) < 1e-14


def test_iso_loop_conserves_mass_and_differs_from_mainloop():
    from oracle import cpuref
    dims = (32, 24)
    n0 = droplet_field(dims, seed=9)
    a, b = cpuref.State('D2Q9', dims), cpuref.State('D2Q9', dims)
    for s in (a, b):
        cpuref.init_pop(s, n0, np.zeros(2 * n0.size))
    m0 = a.pop.sum()
    cpuref.iso_run(a, 'D2E4', G, 1.0, nsteps=200)
    cpuref.meta_run(b, 'D2E4', G, 1.0, nsteps=200)
    assert abs(a.pop.sum() - m0) <= 1e-13 * m0
    assert np.isfinite(a.pop).all()
    assert rel_err(a.pop, b.pop) > 1e-4
