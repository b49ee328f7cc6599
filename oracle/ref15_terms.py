"""Term lists of the Meta class's generated collision and initialisation code
(TEST INFRASTRUCTURE ONLY; never imported by the product package).

The current ShanChenMultiPhase class writes, for every population q, the C
expression

    pop_swap[dst] = <str(expr_q.evalf())>

where expr_q is the sympy expansion of (1 - omega) lpop_q + omega f_eq_q +
F_guo_q with omega = OMEGA substituted first (idpy/LBM/Collisions.py:
251-273, 487-530), and the products n u_a, n u_a u_b, u_a F_b replaced by
symbols whose names are their C products (idpy/LBM/Equilibria.py:240-300,
idpy/LBM/Forcings.py:196-300, pressure_mode 'compute'); K_InitPopulationsMeta
writes str(expand(f_eq_q).evalf()) with the same substitutions
(idpy/LBM/LBMKernelsMeta.py:282-390). A C compiler evaluates each printed
sum left to right and each product left to right.

This module rebuilds those expressions with its own sympy code from the
formulas (f_eq: Equilibria.py:107-138; F_guo: Forcings.py:104-127), prints
them as sympy does, and parses the text back into ordered term lists for
oracle/cpu_ref.cpp: term t of population q is coef_t * v_a * v_b * v_c (up
to three factors, left to right) with variables
    0 ln_0, 1..3 lu_0_0..2, 4..6 d_psi_0..2, 7 lpop_q.
The weights and c2 are the exact rationals of the lattice (the sympy path
evaluates them the same way whether they enter as Rationals or as the
Python floats of idpy/LBM/LBM.py:74-92, up to ties at the 15th digit, which
the tests rule out).
"""
import re
from functools import lru_cache

import numpy as np
import sympy as sp

from .stencils import XI, XI_W

_VARS = {'ln_0': 0, 'lu_0_0': 1, 'lu_0_1': 2, 'lu_0_2': 3,
         'd_psi_0': 4, 'd_psi_1': 5, 'd_psi_2': 6}


def _taylor2(D):
    return [(a, b) for a in range(D) for b in range(a, D)]


def _symbols(D, Q):
    n = sp.Symbol('ln_0')
    u = [sp.Symbol('lu_0_%d' % a) for a in range(D)]
    F = [sp.Symbol('d_psi_%d' % a) for a in range(D)]
    f = [sp.Symbol('lpop_%d' % q) for q in range(Q)]
    return n, u, F, f


def _tuples_eq(n, u, D):
    """(product, symbol named by its C code), higher orders first."""
    out = [(n * u[a], sp.Symbol('ln_0*lu_0_%d' % a)) for a in range(D)]
    out += [(n * u[a] * u[b], sp.Symbol('ln_0*lu_0_%d*lu_0_%d' % (a, b))) for a, b in _taylor2(D)]
    return out[::-1]


def _tuples_guo(u, F, D):
    out = []
    for a, b in _taylor2(D):
        out.append((u[a] * F[b], sp.Symbol('lu_0_%d*d_psi_%d' % (a, b))))
        if a != b:
            out.append((u[b] * F[a], sp.Symbol('lu_0_%d*d_psi_%d' % (b, a))))
    return out[::-1]


def _feq(xi, w, c2, n, u):
    D = len(u)
    h1 = sum(xi[a] * u[a] for a in range(D)) / c2
    h2 = sum((xi[a] * xi[b] - (c2 if a == b else 0)) * u[a] * u[b]
             for a in range(D) for b in range(D)) / (2 * c2 * c2)
    return n * (1 + h1 + h2) * w


def _fguo(xi, w, c2, om, u, F):
    D = len(u)
    g1 = sum(xi[a] * F[a] for a in range(D)) / c2
    g2 = sum((xi[a] * xi[b] - (c2 if a == b else 0)) * (u[a] * F[b] + u[b] * F[a])
             for a in range(D) for b in range(D)) / (2 * c2 * c2)
    return (g1 + g2) * (1 - om / 2) * w


def _parse(text, q):
    """Ordered (coef, [vars]) of a printed sum of 'coef*name*name...' terms."""
    s = text.replace(' ', '')
    terms = []
    for m in re.finditer(r'([+-]?)([^+-]+(?:e[+-]\d+[^+-]*)?)', s):
        sign, body = m.group(1), m.group(2)
        parts = body.split('*')
        coef = float(parts[0])
        if sign == '-':
            coef = -coef
        vs = []
        for p in parts[1:]:
            vs.append(7 if p == 'lpop_%d' % q else _VARS[p])
        terms.append((coef, vs))
    return terms


def printed_collision(xi_name, omega, q):
    """The printed right-hand side of population q of the BGK collision."""
    xis, ws = XI[xi_name], XI_W[xi_name]
    D, Q = len(xis[0]), len(xis)
    c2 = sp.Rational(1, 3)
    n, u, F, f = _symbols(D, Q)
    om = sp.Symbol('omega')
    w = sp.Rational(ws[q].numerator, ws[q].denominator)
    e = (1 - om) * f[q] + om * _feq(xis[q], w, c2, n, u) + _fguo(xis[q], w, c2, om, u, F)
    e = sp.expand(e.subs(om, omega))
    for a, b in _tuples_eq(n, u, D) + _tuples_guo(u, F, D):
        e = e.subs(a, b)
    return str(e.evalf())


def printed_equilibrium(xi_name, q):
    xis, ws = XI[xi_name], XI_W[xi_name]
    D, Q = len(xis[0]), len(xis)
    c2 = sp.Rational(1, 3)
    n, u, F, f = _symbols(D, Q)
    w = sp.Rational(ws[q].numerator, ws[q].denominator)
    e = sp.expand(_feq(xis[q], w, c2, n, u))
    for a, b in _tuples_eq(n, u, D):
        e = e.subs(a, b)
    return str(e.evalf())


def _pack(per_q):
    """Flat arrays for cpu_ref: offsets (Q+1), coefficients (T), factors
    (3 T, -1 for none)."""
    off, coef, fac = [0], [], []
    for terms in per_q:
        for c, vs in terms:
            if not 1 <= len(vs) <= 3:
                raise ValueError('term with %d factors' % len(vs))
            coef.append(c)
            fac.extend(vs + [-1] * (3 - len(vs)))
        off.append(len(coef))
    return (np.array(off, dtype=np.int32), np.array(coef, dtype=np.float64),
            np.array(fac, dtype=np.int32))


@lru_cache(maxsize=None)
def collision_terms(xi_name, omega):
    """Packed term lists of the collision for every population."""
    Q = len(XI[xi_name])
    return _pack([_parse(printed_collision(xi_name, float(omega), q), q) for q in range(Q)])


@lru_cache(maxsize=None)
def equilibrium_terms(xi_name):
    Q = len(XI[xi_name])
    return _pack([_parse(printed_equilibrium(xi_name, q), q) for q in range(Q)])


def monomial_coefficients(xi_name, omega):
    """Per population {monomial key: coefficient} read from the printed text
    (keys as idmx.LBM.ref15: ('f',), ('n',), ('nu', a), ('nuu', a, b),
    ('F', a), ('uF', a, b))."""
    off, coef, fac = collision_terms(xi_name, float(omega))
    out = []
    for q in range(len(off) - 1):
        P = {}
        for t in range(off[q], off[q + 1]):
            vs = [v for v in fac[3 * t:3 * t + 3] if v >= 0]
            us = sorted(v - 1 for v in vs if 1 <= v <= 3)
            Fs = [v - 4 for v in vs if 4 <= v <= 6]
            if 7 in vs:
                k = ('f',)
            elif 0 in vs:
                k = ('n',) if not us else ('nu', us[0]) if len(us) == 1 else ('nuu', us[0], us[1])
            elif not us:
                k = ('F', Fs[0])
            else:
                k = ('uF', us[0], Fs[0])
            P[k] = P.get(k, 0.0) + float(coef[t])
        out.append(P)
    return out
