"""Oracle derivation of the MRT moment bases (TEST INFRASTRUCTURE ONLY).

Never imported by the product package. Used by tests/ and bench.py's
cpu_baseline leg to form the matrices that oracle/cpu_ref.cpp applies.

Follows the reference procedure step by step, in exact rational arithmetic:

* Hermite tensors from derivatives of the Gaussian weight,
  H^(n)_{i..}(c) = (-1)^n d^n w / w            (idpy/Utils/Hermite.py:37-56)
  rescaled as c_s^n H^(n)(xi/c_s), c_s^2 = c2  (IdpyStencils.py:838-866)
* Taylor-tuple order of the tensor components  (idpy/Utils/IdpySymbolic.py:68-88)
* invertible set: walk orders 1, 2, ... keeping rows that are not all zero on
  the lattice until Q + search_depth rows, then keep the rref pivot rows of
  the stacked matrix                           (IdpyStencils.py:905-971)
* W-orthogonal set: classical Gram-Schmidt with <a, b> = sum_i a_i b_i w_i,
  projecting the original row on each earlier orthogonal row
                                               (IdpyStencils.py:999-1043,
                                                idpy/Utils/Geometry.py:97-105,
                                                idpy/Utils/Hermite.py:59-65)
* collision  f* = M^-1 [S M f_eq + (1-S) M f] + M'^-1 (1-S/2) M' F
  with M the W-orthogonal basis and M' the non-orthogonal invertible basis
                                               (Collisions.py:96-160,
                                                Forcings.py:129-181)

The stencil weights and c2 are taken as exact rationals (the values of
IDStencils['LBM']['XI_D2Q9'/'XI_D3Q19'], IdpyStencils.py:60-120). The
Python-float tables of idpy/LBM/LBM.py:74-92 round these to doubles; a basis
built from them is not pinned here.
"""
from functools import lru_cache
from itertools import combinations_with_replacement

import numpy as np
import sympy as sp

XI = {
    'D2Q9': ((0, 0), (1, 0), (0, 1), (-1, 0), (0, -1),
             (1, 1), (-1, 1), (-1, -1), (1, -1)),
    'D3Q19': ((0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0),
              (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0), (-1, -1, 0),
              (1, -1, 0), (1, 0, 1), (-1, 0, 1), (-1, 0, -1), (1, 0, -1),
              (0, 1, 1), (0, -1, 1), (0, -1, -1), (0, 1, -1)),
}
W = {
    'D2Q9': (sp.Rational(4, 9),) + (sp.Rational(1, 9),) * 4 + (sp.Rational(1, 36),) * 4,
    'D3Q19': (sp.Rational(1, 3),) + (sp.Rational(1, 18),) * 6 + (sp.Rational(1, 36),) * 12,
}
C2 = sp.Rational(1, 3)


def taylor_tuples(D, n):
    """Index tuples of a rank-n symmetric tensor in idpy's order
    (IdpySymbolic.py:68-88 equals combinations with replacement)."""
    return list(combinations_with_replacement(range(D), n))


@lru_cache(maxsize=None)
def hermite_poly(D, tup):
    """c_s^n H^(n)_tup(xi / c_s) as a polynomial in xi_0..xi_{D-1}."""
    c = sp.symbols('c_0:%d' % D)
    xi = sp.symbols('xi_0:%d' % D)
    cs = sp.Symbol('c_s', positive=True)
    gauss = sp.exp(-sum(ci ** 2 for ci in c) / 2)
    der = gauss
    for i in tup:
        der = sp.diff(der, c[i])
    h = sp.simplify((-1) ** len(tup) * der / gauss)
    h = h.subs({c[i]: xi[i] / cs for i in range(D)}, simultaneous=True)
    h = sp.expand(sp.simplify(cs ** len(tup) * h))
    return sp.expand(h.subs(cs, sp.sqrt(C2))), xi


def lattice_row(stencil, tup):
    D = len(XI[stencil][0])
    h, xi = hermite_poly(D, tup)
    return [h.subs({xi[d]: v[d] for d in range(D)}) for v in XI[stencil]]


@lru_cache(maxsize=None)
def invertible_set(stencil, search_depth=6):
    """(M, tuples) of GetInvertibleHermiteSet (IdpyStencils.py:905-971)."""
    xis = XI[stencil]
    Q, D = len(xis), len(xis[0])
    rows, tuples = [[sp.Integer(1)] * Q], [()]
    count, order = 1, 1
    while count < Q + search_depth:
        for tup in taylor_tuples(D, order):
            r = lattice_row(stencil, tup)
            if not all(v == 0 for v in r):
                rows.append(r)
                tuples.append(tup)
                count += 1
        order += 1
    full = sp.Matrix(rows)
    _, piv = full.T.rref()
    M = full[list(piv), :]
    return M, [tuples[i] for i in piv]


def _wprod(a, b, w):
    return sum(a[i] * b[i] * w[i] for i in range(len(w)))


@lru_cache(maxsize=None)
def worth_set(stencil, search_depth=6):
    """W-orthogonal basis of GetWOrthInvertibleHermiteSet
    (IdpyStencils.py:999-1043)."""
    M, _ = invertible_set(stencil, search_depth)
    w = W[stencil]
    out = [M[0, :]]
    for i in range(1, M.shape[0]):
        v = M[i, :]
        nv = M[i, :]
        for u in out:
            nv = nv - u * (_wprod(v, u, w) / _wprod(u, u, w))
        out.append(nv)
    return sp.Matrix.vstack(*out)


def mrt_matrices(stencil, S):
    """Dense (A, B, Gg) for oracle/cpu_ref.cpp, with
    A = M^-1 (1-S) M, B = M^-1 S M, Gg = M'^-1 (1-S/2) M'.

    S is a Q x Q matrix in the moment index order of the W-orthogonal basis
    (the reference applies the same S in both bases, Forcings.py:176-177)."""
    Mw = worth_set(stencil)
    Mp, _ = invertible_set(stencil)
    Q = Mw.shape[0]
    S = sp.Matrix(S)
    I = sp.eye(Q)
    A = Mw.inv() * (I - S) * Mw
    B = Mw.inv() * S * Mw
    Gg = Mp.inv() * (I - S / 2) * Mp
    f = lambda X: np.array(X.evalf(30).tolist(), dtype=np.float64)
    return f(A), f(B), f(Gg)


MONOMIALS = {
    'D2Q9': ((0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 1), (2, 1), (1, 2), (2, 2)),
    'D3Q19': ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
              (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2),
              (0, 2, 1), (0, 1, 2), (2, 2, 0), (2, 0, 2), (0, 2, 2)),
}


@lru_cache(maxsize=None)
def moment_tables(stencil):
    """Tables of the moment-space form of the same collision (diagonal S),
    f* = f + W Mm^T Tm^T Dinv (Lt mg - md), from this module's bases:
    Mm[k][q] = xi_q^e_k for the lattice monomials e_k, Tm = Mw Mm^-1,
    Lt = Mw Mp^-1 (unit lower-triangular: Gram-Schmidt), Dinv the inverse
    W-norms of the rows of Mw, the Hermite tuple of every row up to second
    order, the first velocity of every opposite pair and MpV, the Mp
    moments of the basis vectors w o {1, xi_a, xi_a xi_b (a <= b)}."""
    xis = XI[stencil]
    Q, D = len(xis), len(xis[0])
    Mm = sp.Matrix([[sp.prod([sp.Integer(v[a]) ** e[a] for a in range(D)]) for v in xis]
                    for e in MONOMIALS[stencil]])
    Mw = worth_set(stencil)
    Mp, tuples = invertible_set(stencil)
    Tm = Mw * Mm.inv()
    Lt = Mw * Mp.inv()
    w = W[stencil]
    Dg = [_wprod(Mw[k, :], Mw[k, :], w) for k in range(Q)]
    opp = [xis.index(tuple(-c for c in v)) for v in xis]
    pairs = [q for q in range(Q) if q < opp[q]]
    ht = np.zeros((3, Q), dtype=np.int32)
    for k, t in enumerate(tuples):
        ht[0, k] = len(t) if len(t) <= 2 else 3
        ht[1, k] = t[0] if 1 <= len(t) <= 2 else -1
        ht[2, k] = t[1] if len(t) == 2 else -1
    mons = [()] + [(a,) for a in range(D)] + [(a, a) for a in range(D)] + \
        [(a, b) for a in range(D) for b in range(a + 1, D)]
    Vb = sp.Matrix([[w[q] * sp.prod([sp.Integer(xis[q][a]) for a in m]) for q in range(Q)]
                    for m in mons])
    MpV = (Mp * Vb.T).T
    f = lambda X: np.ascontiguousarray(np.array(X.evalf(30).tolist(), dtype=np.float64))
    return {'pairs': np.array(pairs, dtype=np.int32),
            'mm': np.ascontiguousarray(np.array(Mm.tolist(), dtype=np.int32)),
            'odd': np.array([sum(e) % 2 for e in MONOMIALS[stencil]], dtype=np.int32),
            'tm': f(Tm), 'lt': f(Lt),
            'dinv': np.array([float(1 / d) for d in Dg], dtype=np.float64),
            'ht': np.ascontiguousarray(ht), 'mpv': f(MpV)}
