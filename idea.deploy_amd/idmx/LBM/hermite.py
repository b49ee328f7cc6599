"""Hermite moment bases for the MRT collision, in exact rational arithmetic.

The moment basis is the one idpy builds (idpy/IdpyStencils/IdpyStencils.py:
905-1043): tensor Hermite polynomials of the lattice velocities, taken order by
order in Taylor-tuple (combinations-with-replacement) order, dropping rows that
vanish on the lattice, keeping the first Q linearly independent rows (the rref
pivots); the W-orthogonal basis applies Gram-Schmidt with
<a, b> = sum_i a_i b_i w_i to those rows.

Derived here from the closed form of the one-dimensional factors,
h_0 = 1, h_1 = x, h_{k+1}(x) = x h_k(x) - k c2 h_{k-1}(x)
(c_s^k He_k(x/c_s) with c_s^2 = c2), so that a tensor component with index
counts (k_x, k_y, k_z) is h_kx(xi_x) h_ky(xi_y) h_kz(xi_z).

From a relaxation matrix S (in W-orthogonal moment order) the collision
f* = M^-1 [S M f_eq + (1-S) M f] + M'^-1 (1 - S/2) M' F_guo
(idpy/LBM/Collisions.py:96-160, idpy/LBM/Forcings.py:129-181; M' is the
non-orthogonalised basis) is reduced to the three tables the kernel takes:
A = M^-1 (1-S) M, beta_k = M^-1 S M v_k, gamma_k = M'^-1 (1-S/2) M' v_k with
v_k = w * {1, xi_a, xi_a xi_b (a <= b)}.
"""
import math
from fractions import Fraction as Fr
from functools import lru_cache
from itertools import combinations_with_replacement

import numpy as np

from .stencils import XIStencils

C2 = Fr(1, 3)


def _h1d(k, x):
    a, b = Fr(1), Fr(x)
    if k == 0:
        return a
    for j in range(1, k):
        a, b = b, x * b - j * C2 * a
    return b


def _lattice(name):
    st = XIStencils[name]
    xis = st['XIs']
    D = len(xis[0])
    ws = [Fr(w).limit_denominator(1000) for w in st['Ws']]
    return xis, ws, D


def hermite_row(name, tup):
    xis, _, D = _lattice(name)
    counts = [tup.count(a) for a in range(D)]
    row = []
    for v in xis:
        p = Fr(1)
        for a in range(D):
            p *= _h1d(counts[a], v[a])
        row.append(p)
    return row


def _rank_filter(rows):
    """Indices of the first linearly independent rows (greedy, exact)."""
    basis = []   # reduced rows with their pivot column
    keep = []
    for i, r in enumerate(rows):
        v = list(r)
        for piv, b in basis:
            if v[piv] != 0:
                f = v[piv] / b[piv]
                v = [x - f * y for x, y in zip(v, b)]
        nz = [j for j, x in enumerate(v) if x != 0]
        if nz:
            basis.append((nz[0], v))
            keep.append(i)
    return keep


@lru_cache(maxsize=None)
def invertible_set(name, search_depth=6):
    """(rows, tuples): the non-orthogonal invertible Hermite basis M'."""
    xis, _, D = _lattice(name)
    Q = len(xis)
    rows, tuples = [[Fr(1)] * Q], [()]
    order = 1
    while len(rows) < Q + search_depth:
        for tup in combinations_with_replacement(range(D), order):
            r = hermite_row(name, tup)
            if any(x != 0 for x in r):
                rows.append(r)
                tuples.append(tup)
        order += 1
    keep = _rank_filter(rows)
    if len(keep) != Q:
        raise RuntimeError('Hermite basis for %s has rank %d' % (name, len(keep)))
    return tuple(tuple(rows[i]) for i in keep), tuple(tuples[i] for i in keep)


@lru_cache(maxsize=None)
def worth_set(name, search_depth=6):
    """(rows, coeffs): W-orthogonal basis M and, for each row i, the
    Gram-Schmidt coefficients c_ij with M_i = M'_i + sum_j c_ij M_j."""
    rows, _ = invertible_set(name, search_depth)
    _, ws, _ = _lattice(name)
    prod = lambda a, b: sum(x * y * w for x, y, w in zip(a, b, ws))
    out, coeffs = [list(rows[0])], [()]
    for i in range(1, len(rows)):
        v = list(rows[i])
        cs = []
        for u in out:
            c = -prod(rows[i], u) / prod(u, u)
            cs.append(c)
            v = [a + c * b for a, b in zip(v, u)]
        out.append(v)
        coeffs.append(tuple(cs))
    return tuple(tuple(r) for r in out), tuple(coeffs)


def _inv(rows):
    n = len(rows)
    a = [list(r) + [Fr(int(i == j)) for j in range(n)] for i, r in enumerate(rows)]
    for c in range(n):
        p = next(r for r in range(c, n) if a[r][c] != 0)
        a[c], a[p] = a[p], a[c]
        pv = a[c][c]
        a[c] = [x / pv for x in a[c]]
        for r in range(n):
            if r != c and a[r][c] != 0:
                f = a[r][c]
                a[r] = [x - f * y for x, y in zip(a[r], a[c])]
    return [row[n:] for row in a]


@lru_cache(maxsize=None)
def _float_bases(name):
    Mw, _ = worth_set(name)
    Mp, _ = invertible_set(name)
    f = lambda m: np.array([[float(x) for x in r] for r in m], dtype=np.float64)
    return f(Mw), f(_inv(Mw)), f(Mp), f(_inv(Mp))


@lru_cache(maxsize=None)
def iso_filter_tables():
    """(M, Ai) float64 9 x 9 tables of the D2Q9 isotropy filter
    (K_IsotropyFilter, idpy/LBM/LBMKernelsMeta.py:895-1080). M holds the
    moment rows M' (1, x, y, xx, xy, yy, xxy, xyy, xxyy); the filter keeps
    the moments 0-3 and 8 of the normalised populations and sets the raw
    sums of xi_x xi_y, xi_x^2 - xi_y^2, xi_x^2 xi_y and xi_x xi_y^2 to values
    computed from the others, so the new populations are Ai b with Ai the
    inverse of the matrix of those nine rows."""
    Mp, tuples = invertible_set('D2Q9')
    if tuples != ((), (0,), (1,), (0, 0), (0, 1), (1, 1), (0, 0, 1), (0, 1, 1),
                  (0, 0, 1, 1)):
        raise RuntimeError('unexpected D2Q9 Hermite set %r' % (tuples,))
    xis, _, _ = _lattice('D2Q9')
    raw = lambda fn: tuple(Fr(fn(v[0], v[1])) for v in xis)
    rows = list(Mp[:4]) + [raw(lambda x, y: x * y), raw(lambda x, y: x * x - y * y),
                           raw(lambda x, y: x * x * y), raw(lambda x, y: x * y * y), Mp[8]]
    f = lambda m: np.array([[float(x) for x in r] for r in m], dtype=np.float64)
    return f(Mp), f(_inv(rows))


def basis_vectors(name):
    """v_k = w * {1, xi_a, xi_a xi_b (a <= b)} as an NB x Q float64 array, in
    the order the kernel uses: 2D {1, x, y, xx, yy, xy}; 3D {1, x, y, z, xx,
    yy, zz, xy, xz, yz}."""
    st = XIStencils[name]
    xi = np.array(st['XIs'], dtype=np.float64)
    w = np.array(st['Ws'], dtype=np.float64)
    D = xi.shape[1]
    vs = [w.copy()]
    vs += [w * xi[:, a] for a in range(D)]
    vs += [w * xi[:, a] * xi[:, a] for a in range(D)]
    vs += [w * xi[:, a] * xi[:, b] for a in range(D) for b in range(a + 1, D)]
    return np.array(vs)


def mrt_tables(name, S):
    """(A, beta, gamma) float64 tables of the MRT kernel for relaxation
    matrix S (Q x Q, W-orthogonal moment order)."""
    Mw, Mwi, Mp, Mpi = _float_bases(name)
    Q = Mw.shape[0]
    S = np.asarray(S, dtype=np.float64)
    if S.shape != (Q, Q):
        raise ValueError('relaxation matrix must be %d x %d' % (Q, Q))
    I = np.eye(Q)
    A = Mwi @ (I - S) @ Mw
    B = Mwi @ S @ Mw
    G = Mpi @ (I - 0.5 * S) @ Mp
    V = basis_vectors(name)
    beta = (B @ V.T).T
    gamma = (G @ V.T).T
    return (np.ascontiguousarray(A), np.ascontiguousarray(beta),
            np.ascontiguousarray(gamma))


def mrt_kernel_tables(name, S):
    """What SCEngine passes to the kernel for relaxation matrix S: the Q
    rates when S is diagonal (moment-space kernel with compiled-in bases),
    else the dense (A, beta, gamma) tables."""
    S = np.asarray(S, dtype=np.float64)
    if np.count_nonzero(S - np.diag(np.diag(S))) == 0:
        return {'diag': np.ascontiguousarray(np.diag(S))}
    return mrt_tables(name, S)


# Lattice monomials xi_x^a xi_y^b (xi_z^c), exponents <= 2, whose values on
# the velocity set form an invertible Q x Q matrix with entries in {-1, 0, 1}.
MONOMIALS = {
    'D2Q9': ((0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 1), (2, 1), (1, 2), (2, 2)),
    'D3Q19': ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
              (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2),
              (0, 2, 1), (0, 1, 2), (2, 2, 0), (2, 0, 2), (0, 2, 2)),
}


@lru_cache(maxsize=None)
def monomial_tables(name):
    """Exact tables of the moment-space MRT collision.

    With Mm[k][q] = xi_q^e_k (the lattice monomials), the W-orthogonal basis
    is Mw = Tm Mm, the non-orthogonal one Mp = Lt^-1 Mw (Lt unit lower-
    triangular), and Mw W Mw^T = diag(1 / Dinv), so that
        f* = f - Mw^-1 s(Mw f - Mw f_eq) + Mp^-1 (1 - s/2) Mp F
           = f + W Mm^T Tm^T Dinv (Lt mg - md)
    with md = s (Mw f - Mw f_eq) and mg = (1 - s/2) Mp F. Returns the
    integer Mm, the parity of each monomial, Tm, Lt and Dinv as Fractions,
    the Hermite tuple (order, a, b) of every row up to second order, and
    the first velocity of every opposite pair."""
    st = XIStencils[name]
    xis = st['XIs']
    Q, D = len(xis), len(xis[0])
    mons = MONOMIALS[name]
    Mm = [[Fr(math.prod(v[a] ** e[a] for a in range(D))) for v in xis] for e in mons]
    Mw, _ = worth_set(name)
    Mp, tuples = invertible_set(name)
    Mmi = _inv(Mm)
    mul = lambda A, B: [[sum(A[i][k] * B[k][j] for k in range(len(B)))
                         for j in range(len(B[0]))] for i in range(len(A))]
    Tm = mul([list(r) for r in Mw], Mmi)
    Lt = mul([list(r) for r in Mw], _inv([list(r) for r in Mp]))
    for i in range(Q):
        if Lt[i][i] != 1 or any(Lt[i][j] != 0 for j in range(i + 1, Q)):
            raise RuntimeError('Mw is not a unit lower-triangular map of Mp')
    _, ws, _ = _lattice(name)
    Dg = [sum(Mw[k][q] * Mw[k][q] * ws[q] for q in range(Q)) for k in range(Q)]
    for k in range(Q):
        for j in range(k):
            if sum(Mw[k][q] * Mw[j][q] * ws[q] for q in range(Q)) != 0:
                raise RuntimeError('Mw is not W-orthogonal')
    xs = [tuple(v) for v in xis]
    opp = [xs.index(tuple(-c for c in v)) for v in xs]
    pairs = [q for q in range(Q) if q < opp[q]]
    ht_ord, ht_a, ht_b = [], [], []
    for t in tuples:
        ht_ord.append(len(t) if len(t) <= 2 else 3)
        ht_a.append(t[0] if 1 <= len(t) <= 2 else -1)
        ht_b.append(t[1] if len(t) == 2 else -1)
    return {'Mm': [[int(x) for x in r] for r in Mm],
            'odd': [sum(e) % 2 for e in mons], 'Tm': Tm, 'Lt': Lt,
            'Dinv': [1 / d for d in Dg], 'ht_ord': ht_ord, 'ht_a': ht_a, 'ht_b': ht_b,
            'pairs': pairs}


@lru_cache(maxsize=None)
def basis_moments(name):
    """(MwV, MpV): the W-orthogonal and the non-orthogonal moments of the
    basis vectors v_b = w * {1, xi_a, xi_a xi_b (a <= b)} (in the order of
    basis_vectors), as exact Fractions, NB x Q."""
    xis, ws, D = _lattice(name)
    Q = len(xis)
    mons = [()] + [(a,) for a in range(D)] + [(a, a) for a in range(D)] + \
        [(a, b) for a in range(D) for b in range(a + 1, D)]
    V = [[ws[q] * math.prod(Fr(xis[q][a]) for a in m) for q in range(Q)] for m in mons]
    Mw, _ = worth_set(name)
    Mp, _ = invertible_set(name)
    mom = lambda M: [[sum(M[k][q] * v[q] for q in range(Q)) for k in range(Q)] for v in V]
    return mom(Mw), mom(Mp)


@lru_cache(maxsize=None)
def moment_polys(name, worth=True, root_xi_sym='\\xi'):
    """The moment basis as polynomials in the symbols root_xi_sym_0..D-1
    (sympy): the Hermite polynomials c_s^n H^(n)(xi / c_s) of the invertible
    set, or with worth=True their Gram-Schmidt combinations (the same
    coefficients as the rows of the W-orthogonal basis)."""
    import sympy as sp
    _, tuples = invertible_set(name)
    _, _, D = _lattice(name)
    xi = [sp.Symbol('%s_%d' % (root_xi_sym, a)) for a in range(D)]
    c2 = sp.Rational(C2.numerator, C2.denominator)

    def h1(k, x):
        a, b = sp.Integer(1), x
        if k == 0:
            return a
        for j in range(1, k):
            a, b = b, sp.expand(x * b - j * c2 * a)
        return b

    raw = [sp.expand(sp.Mul(*[h1(t.count(a), xi[a]) for a in range(D)])) for t in tuples]
    if not worth:
        return tuple(raw)
    _, coeffs = worth_set(name)
    out = []
    for i, p in enumerate(raw):
        q = p
        for j, c in enumerate(coeffs[i]):
            q = q + sp.Rational(c.numerator, c.denominator) * out[j]
        out.append(sp.expand(q))
    return tuple(out)
