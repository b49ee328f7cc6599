"""Pins of the CPU oracle (no GPU): closed forms, invariants and an
independent numpy brute force, chosen so that a dropped term, a wrong sign
or index, or a transposed operand in oracle/cpu_ref.cpp fails one of them.
The reference holds no golden vectors for this path (idpy/LBM/test.py:84-101
asserts True), so these are the pins."""
import itertools
from fractions import Fraction as Fr

import numpy as np
import pytest

from helpers import droplet_field, rel_err

G = -5.0 / 3.0


# ---------------------------------------------------------------- stencils
def _moment(vs, ws, idx):
    return sum(w * np.prod([v[i] for i in idx]) for v, w in zip(vs, ws))


@pytest.mark.parametrize('name', ['D2Q9', 'D3Q19'])
def test_velocity_set_isotropy(name):
    """sum w = 1, sum w xi_a xi_b = c2 delta, 4th moment = c2^2 (dd+dd+dd),
    odd moments vanish (exact rationals)."""
    from oracle.stencils import XI, XI_W
    vs, ws = XI[name], XI_W[name]
    D = len(vs[0])
    c2 = Fr(1, 3)
    assert sum(ws) == 1
    for a, b in itertools.product(range(D), repeat=2):
        assert _moment(vs, ws, (a, b)) == (c2 if a == b else 0)
        assert _moment(vs, ws, (a,)) == 0
    d = lambda i, j: 1 if i == j else 0
    for a, b, c, e in itertools.product(range(D), repeat=4):
        want = c2 * c2 * (d(a, b) * d(c, e) + d(a, c) * d(b, e) + d(a, e) * d(b, c))
        assert _moment(vs, ws, (a, b, c, e)) == want
        assert _moment(vs, ws, (a, b, c)) == 0


@pytest.mark.parametrize('name,order', [('D2E4', 4), ('D2E6', 6), ('D2E8', 8),
                                        ('D3E4', 4), ('D3E6', 6), ('D3E8', 8)])
def test_estencil_isotropy(name, order):
    """sum w e_a e_b = delta (e2 = 1), and the rank-4 (E4), rank-6 (E6) and
    rank-8 (E8) moment tensors have the component ratios of an isotropic
    tensor (sums over pairings of deltas)."""
    from oracle.stencils import E
    vs, ws = E[name]
    D = len(vs[0])
    for a, b in itertools.product(range(D), repeat=2):
        assert _moment(vs, ws, (a, b)) == (1 if a == b else 0)
    # rank 4 isotropy: M_xxxx = 3 M_xxyy (needed for E4 and up)
    assert _moment(vs, ws, (0, 0, 0, 0)) == 3 * _moment(vs, ws, (0, 0, 1, 1))
    if order >= 6:
        # rank 6: 15 pairings; x^6 : x^4 y^2 : x^2 y^2 z^2 = 15 : 3 : 1
        assert _moment(vs, ws, (0,) * 6) == 5 * _moment(vs, ws, (0, 0, 0, 0, 1, 1))
        if D == 3:
            assert _moment(vs, ws, (0, 0, 0, 0, 1, 1)) == 3 * _moment(vs, ws, (0, 0, 1, 1, 2, 2))
    if order >= 8:
        assert _moment(vs, ws, (0,) * 8) == 7 * _moment(vs, ws, (0,) * 6 + (1, 1))
        # rank 8: x^8 : x^6 y^2 : x^4 y^4 = 105 : 15 : 9
        assert 5 * _moment(vs, ws, (0,) * 4 + (1,) * 4) == 3 * _moment(vs, ws, (0,) * 6 + (1, 1))


def test_product_tables_equal_oracle_tables():
    from idmx.LBM.stencils import FStencils, XIStencils
    from oracle.stencils import E, XI, XI_W
    for k in XI:
        assert tuple(XIStencils[k]['XIs']) == XI[k]
        assert all(a == float(b) for a, b in zip(XIStencils[k]['Ws'], XI_W[k]))
    for k in E:
        assert tuple(FStencils[k]['Es']) == E[k][0]
        assert all(a == float(b) for a, b in zip(FStencils[k]['Ws'], E[k][1]))


# --------------------------------------------------------- force closed form
@pytest.mark.parametrize('dims', [(32, 8), (16, 4, 4)])
def test_sc_force_plane_wave_closed_form(dims):
    """psi = n, n = 1 + a sin(k x): with E4, sum_e w_e e_x psi(x+e) =
    a cos(k x) sin(k) exactly (2/3 sin k from the axis pair, 1/3 sin k from
    the four diagonals), so F_x = -G n(x) a sin(k) cos(k x), F_y = F_z = 0."""
    from oracle import cpuref
    xi = 'D2Q9' if len(dims) == 2 else 'D3Q19'
    es = 'D2E4' if len(dims) == 2 else 'D3E4'
    L0 = dims[0]
    k = 2 * np.pi * 3 / L0
    a = 0.01
    x = np.arange(int(np.prod(dims))) % L0
    n = 1.0 + a * np.sin(k * x)
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n, np.zeros(n.size * len(dims)))
    cpuref.meta_moments(st, es, G, psi='linear')
    V = n.size
    # u_x = F_x / (2 n) at rest populations
    Fx = 2 * n * st.u[:V]
    want = -G * n * a * np.sin(k) * np.cos(k * x)
    assert np.abs(Fx - want).max() < 1e-15
    assert np.abs(st.u[V:]).max() < 1e-16


# ------------------------------------------------------ numpy brute force
@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (16, 12)), ('D2Q9', 'D2E6', (15, 9)),
                                        ('D2Q9', 'D2E8', (13, 11)), ('D3Q19', 'D3E4', (8, 7, 6)),
                                        ('D3Q19', 'D3E6', (7, 6, 5)),
                                        ('D3Q19', 'D3E8', (7, 6, 5))])
def test_oracle_matches_numpy_brute_force(xi, es, dims):
    from oracle import cpuref, np_ref
    n0 = droplet_field(dims, seed=1)
    V, Q = n0.size, (9 if xi == 'D2Q9' else 19)
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0, np.zeros(V * len(dims)))
    f = st.pop.reshape((Q,) + tuple(np.flip(dims))).copy()
    for _ in range(5):
        f, n, psi, u = np_ref.step(f, xi, es, G, 1.0)
    cpuref.meta_run(st, es, G, 1.0, nsteps=5)
    assert rel_err(st.pop, f.ravel()) < 1e-14
    assert rel_err(st.n, n.ravel()) < 1e-14
    # tau != 1 exercises the (1 - omega) f and (1 - omega/2) terms
    st2 = cpuref.State(xi, dims)
    cpuref.init_pop(st2, n0, np.zeros(V * len(dims)))
    f = st2.pop.reshape(f.shape).copy()
    for _ in range(5):
        f, *_ = np_ref.step(f, xi, es, G, 0.8)
    cpuref.meta_run(st2, es, G, 0.8, nsteps=5)
    assert rel_err(st2.pop, f.ravel()) < 1e-14


def test_legacy_pull_equals_meta_push_and_aos_equals_soa():
    """The legacy class pulls (f(x) <- f*(x - xi)), the Meta class pushes
    (f(x + xi) <- f*(x)); both must give the same state, as must AoS and SoA
    orderings."""
    from oracle import cpuref
    dims = (20, 17)
    n0 = droplet_field(dims, seed=2)
    V = n0.size
    a = cpuref.State('D2Q9', dims)
    cpuref.init_pop(a, n0, np.zeros(2 * V))
    b = cpuref.State('D2Q9', dims)
    b.pop[:] = a.pop
    c = cpuref.State('D2Q9', dims, aos=True)
    c.pop[:] = a.pop.reshape(9, V).T.ravel()
    cpuref.legacy_run(a, 'D2E4', G, 1.0, nsteps=25)
    cpuref.meta_run(b, 'D2E4', G, 1.0, nsteps=25)
    cpuref.meta_run(c, 'D2E4', G, 1.0, nsteps=25)
    assert rel_err(a.pop, b.pop) < 1e-15
    assert rel_err(c.pop.reshape(V, 9).T.ravel(), b.pop) < 1e-15


# ---------------------------------------------------------------- invariants
def test_streaming_is_a_spike_permutation():
    """With G = 0, tau = 1 and a single-site spike f_q = w_q at rest, one step
    moves population q by exactly xi_q (the pattern of idpy/LBM/test.py:103)."""
    from oracle import cpuref
    from oracle.stencils import XI
    dims = (9, 7)
    V = 63
    n = np.full(V, 1e-300)
    g0 = 3 + 9 * 4
    n[g0] = 1.0
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n, np.zeros(2 * V))
    before = st.pop.reshape(9, V).copy()
    cpuref.meta_run(st, 'D2E4', 0.0, 1.0, psi='linear', nsteps=1)
    after = st.pop.reshape(9, V)
    for q, (cx, cy) in enumerate(XI['D2Q9']):
        gt = ((3 + cx) % 9) + 9 * ((4 + cy) % 7)
        assert after[q, gt] == pytest.approx(before[q, g0], rel=1e-15)


def test_mass_and_momentum_conserved():
    from oracle import cpuref
    dims = (24, 20)
    n0 = droplet_field(dims, seed=4)
    V = n0.size
    st = cpuref.State('D2Q9', dims)
    cpuref.init_pop(st, n0, np.zeros(2 * V))
    f0 = st.pop.reshape(9, V)
    m0 = f0.sum()
    cpuref.meta_run(st, 'D2E4', G, 1.0, nsteps=300)
    f1 = st.pop.reshape(9, V)
    from oracle.stencils import XI
    xi = np.array(XI['D2Q9'], dtype=float)
    assert abs(f1.sum() - m0) / m0 < 1e-13
    assert np.abs(xi.T @ f1.sum(axis=1)).max() < 1e-12 * m0


def test_equilibrium_and_guo_moments():
    """sum f_eq = n, sum xi f_eq = n u, sum xi xi f_eq = n (c2 I + u u);
    sum F_guo = 0 and sum xi F_guo = F for the raw Guo term."""
    from oracle import np_ref
    from oracle.stencils import XI
    rng = np.random.default_rng(0)
    for name in ('D2Q9', 'D3Q19'):
        xi = np.array(XI[name], dtype=float)
        D = xi.shape[1]
        n = rng.uniform(0.1, 2.0, 5)
        u = [rng.uniform(-0.1, 0.1, 5) for _ in range(D)]
        F = [rng.uniform(-0.1, 0.1, 5) for _ in range(D)]
        feq, fg = np_ref.feq_guo(name, n, u, F)
        assert np.allclose(feq.sum(0), n, atol=1e-15, rtol=0)
        assert np.allclose(fg.sum(0), 0, atol=1e-15)
        for a in range(D):
            assert np.allclose(xi[:, a] @ feq, n * u[a], atol=1e-15)
            assert np.allclose(xi[:, a] @ fg, F[a], atol=1e-15)
            for b in range(D):
                want = n * ((1 / 3 if a == b else 0) + u[a] * u[b])
                assert np.allclose((xi[:, a] * xi[:, b]) @ feq, want, atol=1e-15)


# ------------------------------------------------------------------- MRT
@pytest.mark.parametrize('name', ['D2Q9', 'D3Q19'])
def test_mrt_basis_is_w_orthogonal_hermite(name):
    import sympy as sp
    from oracle import hermite_mrt as hm
    M, tuples = hm.invertible_set(name)
    Mw = hm.worth_set(name)
    W = sp.diag(*hm.W[name])
    Gm = Mw * W * Mw.T
    Q = M.shape[0]
    assert all(Gm[i, j] == 0 for i in range(Q) for j in range(Q) if i != j)
    assert M.rank() == Q
    # row (0, 0) is xi_x^2 - c2; row (0,) is xi_x
    i = tuples.index((0, 0))
    assert [M[i, q] for q in range(Q)] == [sp.Rational(v[0] ** 2) - sp.Rational(1, 3)
                                           for v in hm.XI[name]]
    assert [M[tuples.index((0,)), q] for q in range(Q)] == [v[0] for v in hm.XI[name]]


# Closed-form tensor Hermite polynomials c_s^n H^(n)(xi / c_s), c2 = 1/3, as
# products of the one-dimensional ones (index counts (kx, ky, kz)):
# h0 = 1, h1 = x, h2 = x^2 - c2, h3 = x^3 - 3 c2 x, h4 = x^4 - 6 c2 x^2 + 3 c2^2.
_C2 = Fr(1, 3)
_H1D = (lambda x: Fr(1), lambda x: Fr(x), lambda x: x * x - _C2,
        lambda x: x ** 3 - 3 * _C2 * x, lambda x: x ** 4 - 6 * _C2 * x * x + 3 * _C2 * _C2)


def _hermite_closed(tup, v):
    out = Fr(1)
    for a in range(len(v)):
        out *= _H1D[tup.count(a)](Fr(v[a]))
    return out


# The rows of the non-orthogonal invertible basis M' in the reference's order
# (GetInvertibleHermiteSet, idpy/IdpyStencils/IdpyStencils.py:905-971:
# rank by rank, Taylor-tuple order within a rank, rows that vanish on the
# lattice skipped, then the rref pivot rows), written out by hand.
_MP_ROWS = {
    'D2Q9': [
        ((), lambda x, y: 1),
        ((0,), lambda x, y: x),
        ((1,), lambda x, y: y),
        ((0, 0), lambda x, y: x * x - _C2),
        ((0, 1), lambda x, y: x * y),
        ((1, 1), lambda x, y: y * y - _C2),
        ((0, 0, 1), lambda x, y: x * x * y - _C2 * y),
        ((0, 1, 1), lambda x, y: x * y * y - _C2 * x),
        ((0, 0, 1, 1), lambda x, y: x * x * y * y - _C2 * (x * x + y * y) + _C2 * _C2),
    ],
    'D3Q19': [
        ((), lambda x, y, z: 1),
        ((0,), lambda x, y, z: x),
        ((1,), lambda x, y, z: y),
        ((2,), lambda x, y, z: z),
        ((0, 0), lambda x, y, z: x * x - _C2),
        ((0, 1), lambda x, y, z: x * y),
        ((0, 2), lambda x, y, z: x * z),
        ((1, 1), lambda x, y, z: y * y - _C2),
        ((1, 2), lambda x, y, z: y * z),
        ((2, 2), lambda x, y, z: z * z - _C2),
        ((0, 0, 1), lambda x, y, z: x * x * y - _C2 * y),
        ((0, 0, 2), lambda x, y, z: x * x * z - _C2 * z),
        ((0, 1, 1), lambda x, y, z: x * y * y - _C2 * x),
        ((0, 2, 2), lambda x, y, z: x * z * z - _C2 * x),
        ((1, 1, 2), lambda x, y, z: y * y * z - _C2 * z),
        ((1, 2, 2), lambda x, y, z: y * z * z - _C2 * y),
        ((0, 0, 1, 1), lambda x, y, z: x * x * y * y - _C2 * (x * x + y * y) + _C2 * _C2),
        ((0, 0, 2, 2), lambda x, y, z: x * x * z * z - _C2 * (x * x + z * z) + _C2 * _C2),
        ((1, 1, 2, 2), lambda x, y, z: y * y * z * z - _C2 * (y * y + z * z) + _C2 * _C2),
    ],
}


def _frac_rank(rows):
    import sympy as sp
    return sp.Matrix([[sp.Rational(x.numerator, x.denominator) for x in r] for r in rows]).rank()


@pytest.mark.parametrize('name', ['D2Q9', 'D3Q19'])
def test_mrt_basis_rows_closed_form(name):
    """Every row of M' (oracle and package) equals the hand-written Hermite
    polynomial on the lattice, in the hand-written order; every candidate of
    rank <= 4 left out is zero on the lattice or depends on the rows kept
    before it (the greedy rref pivot rule); the W-orthogonal basis is
    Mw = L M' with L unit lower-triangular and Mw W Mw^T diagonal, which
    fixes the Gram-Schmidt result uniquely."""
    from itertools import combinations_with_replacement
    import sympy as sp
    from idmx.LBM import hermite
    from oracle import hermite_mrt as hm
    xis = hm.XI[name]
    D = len(xis[0])
    want = _MP_ROWS[name]
    Q = len(xis)
    assert len(want) == Q
    rows_cf = [[Fr(fn(*[Fr(c) for c in v])) for v in xis] for _, fn in want]
    # the hand-written polynomials are the product-form Hermite tensors
    for (tup, _), r in zip(want, rows_cf):
        assert r == [_hermite_closed(tup, v) for v in xis]
    # oracle basis
    Mp, tuples = hm.invertible_set(name)
    assert [tuple(t) for t in tuples] == [t for t, _ in want]
    assert all(Mp[i, q] == sp.Rational(rows_cf[i][q].numerator, rows_cf[i][q].denominator)
               for i in range(Q) for q in range(Q))
    # package basis (its own velocity order)
    pk_rows, pk_tuples = hermite.invertible_set(name)
    from idmx.LBM.stencils import XIStencils
    pk_xis = XIStencils[name]['XIs']
    assert list(pk_tuples) == [t for t, _ in want]
    for (tup, fn), r in zip(want, pk_rows):
        assert list(r) == [Fr(fn(*[Fr(c) for c in v])) for v in pk_xis]
    # selection: a candidate not kept is zero on the lattice or dependent on
    # the rows kept before it in Taylor-tuple order
    kept = [t for t, _ in want]
    for order in range(1, 5):
        for tup in combinations_with_replacement(range(D), order):
            if tup in kept:
                continue
            r = [_hermite_closed(tup, v) for v in xis]
            if all(x == 0 for x in r):
                continue
            before = [rows_cf[i] for i, t in enumerate(kept)
                      if (len(t), t) < (len(tup), tup)]
            assert _frac_rank(before + [r]) == _frac_rank(before), tup
    # Gram-Schmidt: Mw = L M', L unit lower-triangular, Mw W Mw^T diagonal
    rat = lambda rows: sp.Matrix([[sp.Rational(x.numerator, x.denominator) for x in r]
                                  for r in rows])
    pk_W = sp.diag(*[sp.Rational(w).limit_denominator(1000) for w in XIStencils[name]['Ws']])
    for Mw, Mpb, W in ((hm.worth_set(name), Mp, sp.diag(*hm.W[name])),
                       (rat(hermite.worth_set(name)[0]), rat(pk_rows), pk_W)):
        L = Mw * Mpb.inv()
        assert all(L[i, i] == 1 for i in range(Q))
        assert all(L[i, j] == 0 for i in range(Q) for j in range(i + 1, Q))
        Gm = Mw * W * Mw.T
        assert all(Gm[i, j] == 0 for i in range(Q) for j in range(Q) if i != j)
        assert all(Gm[i, i] > 0 for i in range(Q))


def test_mrt_with_uniform_rates_is_bgk():
    """S = omega I: A = (1-omega) I, B = omega I, Gg = (1-omega/2) I, so the
    MRT step equals the BGK step (Collisions.py:96-160 vs 163-185)."""
    import sympy as sp
    from oracle import cpuref, hermite_mrt as hm
    dims = (14, 12)
    n0 = droplet_field(dims, seed=6)
    V = n0.size
    tau = 0.8
    mats = hm.mrt_matrices('D2Q9', sp.eye(9) / sp.Rational(4, 5))
    a = cpuref.State('D2Q9', dims)
    cpuref.init_pop(a, n0, np.zeros(2 * V))
    b = cpuref.State('D2Q9', dims)
    b.pop[:] = a.pop
    cpuref.meta_run(a, 'D2E4', G, tau, nsteps=10)
    cpuref.meta_run(b, 'D2E4', G, tau, nsteps=10, mrt=mats)
    assert rel_err(b.pop, a.pop) < 1e-14


@pytest.mark.parametrize('name,es,dims', [('D2Q9', 'D2E4', (24, 21)),
                                           ('D3Q19', 'D3E4', (11, 10, 9))])
def test_mrt_moment_space_equals_dense(name, es, dims):
    """The moment-space evaluation of the MRT collision (lattice monomials,
    opposite-pair sums, Mw = Tm Mm, Mw = Lt Mp, Mw^-1 = W Mw^T Dinv) gives
    the dense f* = A f + B f_eq + Gg F of the same diagonal S of distinct
    rates to rounding over 30 steps. Planted errors fail: in D3Q19 the Guo
    part without the Lt mixing (the forcing taken in the W-orthogonal basis:
    for D2Q9 Lt = I, the two bases coincide), in D2Q9 a dropped c2^2 term of
    the xxyy moment."""
    import sympy as sp
    from oracle import cpuref, hermite_mrt
    Q = 9 if name == 'D2Q9' else 19
    s = np.random.default_rng(3).uniform(0.6, 1.6, size=Q)
    n0 = droplet_field(dims, seed=11)
    V = n0.size
    a = cpuref.State(name, dims)
    cpuref.init_pop(a, n0, np.zeros(V * len(dims)))
    b = cpuref.State(name, dims)
    b.pop[:] = a.pop
    c = cpuref.State(name, dims)
    c.pop[:] = a.pop
    S = sp.diag(*[sp.Rational(v).limit_denominator(10 ** 15) for v in s])
    cpuref.meta_run(a, es, G, 1.0, nsteps=30, mrt=hermite_mrt.mrt_matrices(name, S))
    cpuref.meta_run_mrt_moment(b, es, G, s, nsteps=30)
    assert rel_err(b.pop, a.pop) < 1e-13
    # the tables themselves: Tm Mm = Mw, Lt Mp = Mw, W-orthogonality
    t = hermite_mrt.moment_tables(name)
    Mw = np.array(hermite_mrt.worth_set(name).evalf(30).tolist(), dtype=float)
    Mp = np.array(hermite_mrt.invertible_set(name)[0].evalf(30).tolist(), dtype=float)
    assert np.abs(t['tm'] @ t['mm'] - Mw).max() < 1e-14
    assert np.abs(t['lt'] @ Mp - Mw).max() < 1e-14
    w = np.array([float(x) for x in hermite_mrt.W[name]])
    assert np.abs(Mw @ np.diag(w) @ Mw.T @ np.diag(t['dinv']) - np.eye(Q)).max() < 1e-13
    orig = hermite_mrt.moment_tables.__wrapped__(name)
    if name == 'D3Q19':
        assert np.abs(t['lt'] - np.eye(Q)).max() > 0.1
        planted = dict(orig, lt=np.ascontiguousarray(np.eye(Q)))
    else:
        assert np.abs(t['lt'] - np.eye(Q)).max() == 0.0
        tm = orig['tm'].copy()
        tm[8, 0] = 0.0
        planted = dict(orig, tm=np.ascontiguousarray(tm))
    import oracle.hermite_mrt as hm
    saved = hm.moment_tables
    hm.moment_tables = lambda n: planted
    try:
        cpuref.meta_run_mrt_moment(c, es, G, s, nsteps=30)
    finally:
        hm.moment_tables = saved
    assert rel_err(c.pop, a.pop) > 1e-9


def test_mrt_product_basis_equals_oracle_basis():
    """The package derives the same bases from the 1D Hermite recurrence in
    Fractions; the oracle from derivatives of the Gaussian in sympy."""
    import sympy as sp
    from idmx.LBM import hermite
    from oracle import hermite_mrt as hm
    for name in ('D2Q9', 'D3Q19'):
        rows, tups = hermite.invertible_set(name)
        M, t2 = hm.invertible_set(name)
        assert tups == tuple(tuple(t) for t in t2)
        assert all(sp.Rational(rows[i][j].numerator, rows[i][j].denominator) == M[i, j]
                   for i in range(len(rows)) for j in range(len(rows)))
        w, _ = hermite.worth_set(name)
        Mw = hm.worth_set(name)
        assert all(sp.Rational(w[i][j].numerator, w[i][j].denominator) == Mw[i, j]
                   for i in range(len(w)) for j in range(len(w)))


def test_mrt_tables_uniform_rates():
    from idmx.LBM import hermite
    for name, Q in (('D2Q9', 9), ('D3Q19', 19)):
        om = 1.3
        A, beta, gamma = hermite.mrt_tables(name, om * np.eye(Q))
        V = hermite.basis_vectors(name)
        assert np.abs(A - (1 - om) * np.eye(Q)).max() < 1e-14
        assert np.abs(beta - om * V).max() < 1e-14
        assert np.abs(gamma - (1 - om / 2) * V).max() < 1e-14
