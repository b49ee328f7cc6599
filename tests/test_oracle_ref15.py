"""The current class's printed 15-digit coefficients ("ref15"): the oracle's
term lists against the closed-form coefficients rounded to 15 significant
digits, the product's dense tables against the printed coefficients, and the
analytic mass drift those coefficients imply against the oracle's run."""
from decimal import Decimal, getcontext
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import cpuref, ref15_terms
from oracle.stencils import XI, XI_W

G = -5.0 / 3.0


def _round15(x):
    """The double nearest to x written with 15 significant digits."""
    getcontext().prec = 15
    return float(+Decimal(x.numerator) / Decimal(x.denominator))


def _closed_form(xi_name, omega):
    """Exact coefficient of every monomial of the BGK + Guo collision with
    tau = 1/omega (omega a Fraction): f_eq = w n (1 + 3 xi.u + 4.5 (xi.u)^2 -
    1.5 u.u), F_guo = (1 - omega/2) w (3 (xi - u).F + 9 (xi.u)(xi.F))."""
    out = []
    for xi, w in zip(XI[xi_name], XI_W[xi_name]):
        D = len(xi)
        P = {('n',): omega * w}
        if omega != 1:
            P[('f',)] = 1 - omega
        g = (1 - omega / 2) * w
        for a in range(D):
            P[('nu', a)] = omega * w * 3 * xi[a]
            P[('F', a)] = g * 3 * xi[a]
            for b in range(D):
                P[('uF', a, b)] = g * (9 * xi[a] * xi[b] - (3 if a == b else 0))
                if a < b:
                    P[('nuu', a, b)] = omega * w * 9 * xi[a] * xi[b]
            P[('nuu', a, a)] = omega * w * (Fr(9, 2) * xi[a] * xi[a] - Fr(3, 2))
        out.append({k: v for k, v in P.items() if v != 0})
    return out


@pytest.mark.parametrize('xi_name', ['D2Q9', 'D3Q19'])
@pytest.mark.parametrize('tau', [Fr(1), Fr(4, 5), Fr(5, 4)])
def test_printed_terms_are_closed_forms_rounded_to_15_digits(xi_name, tau):
    omega = 1 / tau
    want = _closed_form(xi_name, omega)
    got = ref15_terms.monomial_coefficients(xi_name, float(omega))
    for q, (P, W) in enumerate(zip(got, want)):
        assert set(P) == set(W), (q, sorted(set(P) ^ set(W)))
        for k, v in W.items():
            assert P[k] == _round15(v), (q, k, P[k], float(v))


@pytest.mark.parametrize('xi_name', ['D2Q9', 'D3Q19'])
def test_printed_equilibrium_terms(xi_name):
    off, coef, fac = ref15_terms.equilibrium_terms(xi_name)
    want = _closed_form(xi_name, Fr(1))
    for q in range(len(off) - 1):
        n_terms = {tuple(v for v in fac[3 * t:3 * t + 3] if v >= 0): coef[t]
                   for t in range(off[q], off[q + 1])}
        # the equilibrium holds exactly the n-monomials of the omega = 1 collision
        eq = {k: v for k, v in want[q].items() if k[0] in ('n', 'nu', 'nuu')}
        assert len(n_terms) == len(eq)
        assert n_terms[(0,)] == _round15(eq[('n',)])


@pytest.mark.parametrize('xi_name', ['D2Q9', 'D3Q19'])
@pytest.mark.parametrize('tau', [1.0, 0.8, 1.25])
def test_product_tables_hold_the_printed_coefficients(xi_name, tau):
    """The dense tables the library steps with multiply each monomial by its
    printed coefficient to within a few units in the last place (the
    printed rounding is ~1e-15 relative)."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'idea.deploy_amd'))
    from idmx.LBM import ref15
    A, beta, gamma = ref15.bgk_tables(xi_name, tau)
    printed = ref15_terms.monomial_coefficients(xi_name, 1.0 / tau)
    D = len(XI[xi_name][0])
    pairs = [(a, a) for a in range(D)] + [(a, b) for a in range(D) for b in range(a + 1, D)]
    for q, P in enumerate(printed):
        b, g = beta[:, q], gamma[:, q]
        M = {('f',): A[q, q], ('n',): b[0]}
        for a in range(D):
            M[('nu', a)] = 3.0 * b[1 + a]
            M[('F', a)] = 3.0 * g[1 + a]
        for k, (a, c) in enumerate(pairs):
            if a == c:
                M[('nuu', a, a)] = -1.5 * b[0] + 4.5 * b[1 + D + k]
                M[('uF', a, a)] = -3.0 * g[0] + 9.0 * g[1 + D + k]
            else:
                M[('nuu', a, c)] = 9.0 * b[1 + D + k]
                M[('uF', a, c)] = M[('uF', c, a)] = 9.0 * g[1 + D + k]
        for k, v in M.items():
            assert abs(v - P.get(k, 0.0)) <= 4e-16 * max(abs(v), 1e-300) + 1e-30, (q, k)
        assert np.all(A[q, np.arange(A.shape[0]) != q] == 0)
    T = ref15.init_table(xi_name)
    eq = ref15_terms.monomial_coefficients(xi_name, 1.0)
    assert np.array_equal(T[0], np.array([P[('n',)] for P in eq]))


def _droplet(dims, seed=3):
    rng = np.random.default_rng(seed)
    ax = np.meshgrid(*[np.arange(L) for L in dims[::-1]], indexing='ij')
    r = np.sqrt(sum((a - L / 2 + 0.3) ** 2 for a, L in zip(ax, dims[::-1])))
    n = 1.05 - 0.95 * np.tanh((r - min(dims) / 4) / 2.0)
    return (n + rng.uniform(-1e-3, 1e-3, size=n.shape)).ravel() * 0.5 + 0.05


def _weight_drift(xi_name, printed):
    """sum_q w_q - 1 in exact arithmetic on the doubles the step multiplies n
    by (tau = 1): the printed 15-digit weights, or the doubles nearest to
    the rationals (the legacy class's W_list, and the oracle's full-double
    step). The relative mass gained per step."""
    conv = _round15 if printed else float
    return float(sum(Fr(conv(w)) for w in XI_W[xi_name]) - 1)


@pytest.mark.parametrize('xi_name,es,dims,nsteps', [('D2Q9', 'D2E4', (32, 32), 2000),
                                                    ('D3Q19', 'D3E4', (12, 12, 10), 1500)])
def test_ref15_mass_drift_is_the_printed_weight_sum(xi_name, es, dims, nsteps):
    """With tau = 1 the printed step multiplies the mass by
    sum_q printed(w_q) every step (the u and F monomials cancel in the
    population sum up to 1e-15 of u.u and u.F), the full-double step by the
    sum of the double weights, 1 - 5.6e-17. Non-uniform fields, so that
    rounding does not add a coherent bias."""
    V = int(np.prod(dims))
    n0 = _droplet(dims)
    delta = _weight_drift(xi_name, True)
    delta_ex = _weight_drift(xi_name, False)
    st = cpuref.State(xi_name, dims)
    cpuref.init_pop_ref15(st, n0, np.zeros(V * len(dims)))
    m_init = float(np.sum(st.pop, dtype=np.longdouble))
    assert abs(m_init / float(np.sum(n0, dtype=np.longdouble)) - 1 - delta) <= 2e-16
    cpuref.meta_run_ref15(st, es, G, 1.0, nsteps=nsteps)
    drift = float(np.sum(st.pop, dtype=np.longdouble)) / m_init - 1
    ex = cpuref.State(xi_name, dims)
    cpuref.init_pop(ex, n0, np.zeros(V * len(dims)))
    m_ex = float(np.sum(ex.pop, dtype=np.longdouble))
    cpuref.meta_run(ex, es, G, 1.0, nsteps=nsteps)
    drift_ex = float(np.sum(ex.pop, dtype=np.longdouble)) / m_ex - 1
    print(xi_name, 'delta', delta, 'drift', drift, 'predicted', nsteps * delta,
          'full-double drift', drift_ex, 'predicted', nsteps * delta_ex)
    # both runs share the rounding of the dynamics (a few 1e-14 here); their
    # difference is the constants' alone
    assert abs(drift_ex - nsteps * delta_ex) <= 5e-14
    want = nsteps * (delta - delta_ex)
    assert abs((drift - drift_ex) - want) <= 0.03 * abs(want)
    assert abs(want) >= 3e-13  # far above the tolerances


def test_ref15_and_exact_agree_to_the_constant_rounding():
    """The printed step and the full-double step on the same input differ by
    the rounding of the constants only: ~1e-15 per step, growing linearly."""
    dims = (24, 20)
    n0 = _droplet(dims, seed=5)
    a, b = cpuref.State('D2Q9', dims), cpuref.State('D2Q9', dims)
    cpuref.init_pop_ref15(a, n0, np.zeros(2 * a.V))
    cpuref.init_pop(b, n0, np.zeros(2 * b.V))
    cpuref.meta_run_ref15(a, 'D2E4', G, 1.0, nsteps=200)
    cpuref.meta_run(b, 'D2E4', G, 1.0, nsteps=200)
    d = np.abs(a.pop - b.pop).max() / np.abs(b.pop).max()
    assert 1e-15 <= d <= 5e-13
