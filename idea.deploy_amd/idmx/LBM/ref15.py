"""The printed constants of the current ShanChenMultiPhase class ("ref15").

idpy's Meta kernels are generated with sympy: the collision of population q
is sp.expand((1 - omega) f_q + omega f_eq_q + F_guo_q) with omega = 1/tau
substituted before the expansion (idpy/LBM/Collisions.py:163-185, 251-273),
the initial populations are sp.expand(f_eq_q) (idpy/LBM/Equilibria.py:
202-210), and each coefficient of the expanded polynomial reaches the C
source through evalf() and str(), i.e. as a decimal of 15 significant digits
(idpy/IdpyCode/IdpyUnroll._codify_sympy). The weights and c2 are the Python
floats of idpy/LBM/LBM.py:74-92. For D2Q9 the printed weights sum to
1 - 8e-16, so that class loses that fraction of its mass at every step with
tau = 1, where the full-double arithmetic of the legacy class conserves it.

This module re-derives the printed coefficients with its own sympy
expansion of the same formulas (f_eq: idpy/LBM/Equilibria.py:107-138; Guo
term: idpy/LBM/Forcings.py:104-127) and maps them onto the tables the
library's dense collision (idmx_sc_desc mrt_A / mrt_beta / mrt_gamma) and
table-driven initialisation (idmx_sc_init_pop_coeffs) take. Those tables
multiply the coordinates c_k, d_k of f_eq and of the raw Guo term on the
basis w o {1, xi_a, xi_a xi_b}; the coordinates are combinations of the
reference's monomials,
  c_0 = n - 1.5 n u.u, c_a = 3 n u_a, c_aa = 4.5 n u_a^2, c_ab = 9 n u_a u_b,
  d_0 = -3 u.F,        d_a = 3 F_a,   d_aa = 9 u_a F_a,   d_ab = 9 (u_a F_b + u_b F_a),
so every table entry is solved from the printed coefficient of one monomial:
the device step then multiplies each monomial by its printed coefficient to
within a few units in the last place, where the printed rounding is a
thousand times larger.

Monomial keys: ('f',) the site's own population, ('n',), ('nu', a),
('nuu', a, b) with a <= b, ('F', a), ('uF', a, b) for u_a F_b.
"""
from functools import lru_cache

import numpy as np
import sympy as sp

from .stencils import XIStencils


def printed(c):
    """The double a C compiler reads from sympy's printed form of the
    number c (evalf() then str(): 15 significant digits)."""
    return float(str(sp.sympify(c).evalf()))


def _syms(D):
    n = sp.Symbol('n')
    u = [sp.Symbol('u_%d' % a) for a in range(D)]
    F = [sp.Symbol('F_%d' % a) for a in range(D)]
    f = sp.Symbol('f')
    return n, u, F, f


def _key(mono, n, u, F, f):
    """Monomial key of a sympy monomial (product of n, u_a, F_a, f)."""
    pw = mono.as_powers_dict()
    if f in pw:
        return ('f',)
    us = sorted(a for a, s in enumerate(u) for _ in range(int(pw.get(s, 0))))
    fs = [a for a, s in enumerate(F) for _ in range(int(pw.get(s, 0)))]
    if n in pw:
        return ('n',) if not us else ('nu', us[0]) if len(us) == 1 else ('nuu', us[0], us[1])
    if len(us) == 0:
        return ('F', fs[0])
    return ('uF', us[0], fs[0])


def _equilibrium(xi, w, c2, n, u):
    """w n (1 + xi.u / c2 + (xi xi - c2 I):u u / (2 c2^2)), the second-order
    Hermite equilibrium (idpy/LBM/Equilibria.py:107-138)."""
    D = len(u)
    h1 = sum(xi[a] * u[a] for a in range(D)) / c2
    h2 = sum((xi[a] * xi[b] - (c2 if a == b else 0)) * u[a] * u[b]
             for a in range(D) for b in range(D)) / (2 * c2 * c2)
    return w * (n * (1 + h1 + h2))


def _guo(xi, w, c2, omega, u, F):
    """(1 - omega/2) w (xi.F / c2 + (xi xi - c2 I):(u F + F u) / (2 c2^2)),
    Guo's term (idpy/LBM/Forcings.py:104-127)."""
    D = len(u)
    g1 = sum(xi[a] * F[a] for a in range(D)) / c2
    g2 = sum((xi[a] * xi[b] - (c2 if a == b else 0)) * (u[a] * F[b] + u[b] * F[a])
             for a in range(D) for b in range(D)) / (2 * c2 * c2)
    return (1 - omega / 2) * w * (g1 + g2)


def _stencil(xi_name):
    st = XIStencils[xi_name]
    return st['XIs'], [sp.Float(w) for w in st['Ws']], sp.Float(st['c2'])


@lru_cache(maxsize=None)
def collision_coefficients(xi_name, tau):
    """Per population, {monomial key: printed coefficient} of the Meta
    class's BGK collision with Guo forcing at relaxation time tau."""
    xis, ws, c2 = _stencil(xi_name)
    D = len(xis[0])
    n, u, F, f = _syms(D)
    om = sp.Symbol('omega')
    out = []
    for q, xi in enumerate(xis):
        e = (1 - om) * f + om * _equilibrium(xi, ws[q], c2, n, u) + _guo(xi, ws[q], c2, om, u, F)
        e = sp.expand(e.subs(om, sp.Float(1.0 / float(tau))))
        out.append({_key(m, n, u, F, f): printed(c) for m, c in e.as_coefficients_dict().items()})
    return tuple(out)


@lru_cache(maxsize=None)
def equilibrium_coefficients(xi_name):
    """Per population, {monomial key: printed coefficient} of the Meta
    class's initial populations f_eq(n, u) (K_InitPopulationsMeta,
    idpy/LBM/LBMKernelsMeta.py:282-390)."""
    xis, ws, c2 = _stencil(xi_name)
    D = len(xis[0])
    n, u, F, f = _syms(D)
    out = []
    for q, xi in enumerate(xis):
        e = sp.expand(_equilibrium(xi, ws[q], c2, n, u))
        out.append({_key(m, n, u, F, f): printed(c) for m, c in e.as_coefficients_dict().items()})
    return tuple(out)


def _pairs(D):
    """Basis order of the tables: 2D {1, x, y, xx, yy, xy}; 3D {1, x, y, z,
    xx, yy, zz, xy, xz, yz}."""
    return [(a, a) for a in range(D)] + [(a, b) for a in range(D) for b in range(a + 1, D)]


def _beta_from(P, D):
    """Equilibrium-part table column (NB values) reproducing the printed
    monomial coefficients P of one population."""
    b0 = P.get(('n',), 0.0)
    col = [b0] + [P.get(('nu', a), 0.0) / 3.0 for a in range(D)]
    for a, b in _pairs(D):
        if a == b:
            col.append((P.get(('nuu', a, a), 0.0) + 1.5 * b0) / 4.5)
        else:
            col.append(P.get(('nuu', a, b), 0.0) / 9.0)
    return col


def _gamma_from(P, D, g0):
    """Guo-part table column: d_0's entry g0 is free (it only enters with the
    diagonal u_a F_a monomials) and kept at its exact value."""
    col = [g0] + [P.get(('F', a), 0.0) / 3.0 for a in range(D)]
    for a, b in _pairs(D):
        if a == b:
            col.append((P.get(('uF', a, a), 0.0) + 3.0 * g0) / 9.0)
        else:
            if P.get(('uF', a, b), 0.0) != P.get(('uF', b, a), 0.0):
                raise RuntimeError('asymmetric u F coefficients')
            col.append(P.get(('uF', a, b), 0.0) / 9.0)
    return col


def bgk_tables(xi_name, tau):
    """(A, beta, gamma) for the dense collision that steps as the Meta
    class's BGK + Guo kernel with its printed coefficients."""
    coef = collision_coefficients(xi_name, float(tau))
    st = XIStencils[xi_name]
    D, Q = len(st['XIs'][0]), len(st['XIs'])
    A = np.zeros((Q, Q))
    beta, gamma = [], []
    om = 1.0 / float(tau)
    for q in range(Q):
        P = coef[q]
        A[q, q] = P.get(('f',), 0.0)
        beta.append(_beta_from(P, D))
        gamma.append(_gamma_from(P, D, (1.0 - 0.5 * om) * st['Ws'][q]))
    return (np.ascontiguousarray(A), np.ascontiguousarray(np.array(beta).T),
            np.ascontiguousarray(np.array(gamma).T))


def init_table(xi_name):
    """NB x Q table for idmx_sc_init_pop_coeffs: f_q = sum_k T[k][q] c_k with
    the printed coefficients of the Meta class's f_eq."""
    coef = equilibrium_coefficients(xi_name)
    D = len(XIStencils[xi_name]['XIs'][0])
    return np.ascontiguousarray(np.array([_beta_from(P, D) for P in coef]).T)


def mrt_tables_printed(xi_name, tables):
    """Dense MRT tables (A, beta, gamma) whose monomial coefficients are
    rounded as the Meta class prints those of its MRT collision
    (CodifySingleMRTCollisionPlusGuoSym, idpy/LBM/Collisions.py:318-374):
    every coefficient of f_j, of the equilibrium monomials and of the Guo
    monomials to 15 significant digits."""
    A, beta, gamma = (np.asarray(t, dtype=np.float64) for t in tables)
    D = 2 if A.shape[0] == 9 else 3
    r = lambda x: float('%.15g' % x)
    A15 = np.vectorize(r)(A)
    B15, G15 = np.zeros_like(beta), np.zeros_like(gamma)
    pairs = _pairs(D)
    for q in range(A.shape[0]):
        b, g = beta[:, q], gamma[:, q]
        P = {('n',): b[0]}
        for a in range(D):
            P[('nu', a)] = 3.0 * b[1 + a]
            P[('F', a)] = 3.0 * g[1 + a]
        for k, (a, c) in enumerate(pairs):
            if a == c:
                P[('nuu', a, a)] = -1.5 * b[0] + 4.5 * b[1 + D + k]
                P[('uF', a, a)] = -3.0 * g[0] + 9.0 * g[1 + D + k]
            else:
                P[('nuu', a, c)] = 9.0 * b[1 + D + k]
                P[('uF', a, c)] = P[('uF', c, a)] = 9.0 * g[1 + D + k]
        P = {k: r(v) for k, v in P.items()}
        B15[:, q] = _beta_from(P, D)
        G15[:, q] = _gamma_from(P, D, g[0])
    return (np.ascontiguousarray(A15), np.ascontiguousarray(B15), np.ascontiguousarray(G15))
