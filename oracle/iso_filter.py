"""Oracle derivation of the D2Q9 isotropy filter (TEST INFRASTRUCTURE ONLY).

Never imported by the product package. Used by tests/ to form the two
9 x 9 tables that oracle/cpu_ref.cpp's filter applies, and to evaluate the
filter the way the reference builds it, as a symbolic solve.

K_IsotropyFilter (idpy/LBM/LBMKernelsMeta.py:895-1080), applied by
MainLoopIsoFilter (idpy/LBM/MultiPhase.py:1099-1223) to the post-collision
populations of every site before the push stream:

* m_k = sum_q M[k][q] f_q with M the invertible Hermite set
  (GetInvertibleHermiteSet, IdpyStencils.py:905-971; rows 1, x, y, xx, xy,
  yy, xxy, xyy, xxyy for D2Q9), and m_k /= m_0 for k >= 1
  (LBMKernelsMeta.py:934-956);
* new populations solve (LBMKernelsMeta.py:962-1004)
      M[k] . f = m_k                 k = 0..3 and 8 (m_0 = 1)
      sum xi_x xi_y f         = m1 m2
      sum (xi_x^2 - xi_y^2) f = m1^2 - m2^2
      sum xi_x^2 xi_y f       = 2 m4 m1 + (m3 + 1/3) m2 - 2 m2 m1^2
      sum xi_x xi_y^2 f       = 2 m4 m2 + (m5 + 1/3) m1 - 2 m1 m2^2
  (the substitutions of LBMKernelsMeta.py:1003-1015);
* the written populations are the solution times m_0 (:1062-1072).

The system is linear in f, so its solution is A^-1 b with A the 9 x 9 matrix
of the left-hand sides and b the right-hand sides; the reference expands the
symbolic solution into monomials of the m_k. The tables are exact rationals
rounded once to doubles.
"""
from functools import lru_cache

import numpy as np
import sympy as sp

from .hermite_mrt import XI, invertible_set

THIRD = sp.Rational(1, 3)


def _row(poly):
    x, y = sp.symbols('x y')
    return [poly(x, y).subs({x: v[0], y: v[1]}) for v in XI['D2Q9']]


@lru_cache(maxsize=None)
def exact_tables():
    """(M, A, Ainv) as sympy matrices: M the moment rows, A the rows of the
    filter's linear system in the order of b above."""
    M, tuples = invertible_set('D2Q9')
    assert tuples == [(), (0,), (1,), (0, 0), (0, 1), (1, 1), (0, 0, 1), (0, 1, 1),
                      (0, 0, 1, 1)], tuples
    rows = [list(M[k, :]) for k in range(4)]
    rows.append(_row(lambda x, y: x * y))
    rows.append(_row(lambda x, y: x ** 2 - y ** 2))
    rows.append(_row(lambda x, y: x ** 2 * y))
    rows.append(_row(lambda x, y: x * y ** 2))
    rows.append(list(M[8, :]))
    A = sp.Matrix(rows)
    return M, A, A.inv()


def tables():
    """(M, Ainv) as float64 9 x 9 arrays for oracle/cpu_ref.cpp."""
    M, _, Ai = exact_tables()
    f = lambda X: np.array(X.evalf(30).tolist(), dtype=np.float64)
    return f(M), f(Ai)


@lru_cache(maxsize=None)
def symbolic_solution():
    """The reference's construction: sympy solve of the system for f with
    the replaced moments as symbols, then the substitutions; returns a
    function of (m1..m8) giving the nine normalised populations, and the
    order of the solved symbols."""
    M, _, _ = exact_tables()
    f = sp.Matrix(sp.symbols('f_0:9'))
    m = sp.Matrix(sp.symbols('m_0:9'))
    mi = sp.Matrix(sp.symbols('mi_0:9'))
    x = [v[0] for v in XI['D2Q9']]
    y = [v[1] for v in XI['D2Q9']]
    dot = lambda c: sum(c[q] * f[q] for q in range(9))
    MH = M * f - m
    eqs = list(MH[:4]) + [
        dot([x[q] * y[q] for q in range(9)]) - mi[4],
        dot([x[q] ** 2 - y[q] ** 2 for q in range(9)]) - mi[5],
        dot([x[q] ** 2 * y[q] for q in range(9)]) - mi[6],
        dot([x[q] * y[q] ** 2 for q in range(9)]) - mi[7],
    ] + list(MH[8:])
    sol = sp.solve(eqs, list(f))
    order = list(sol.keys())
    vec = sp.Matrix([sol[k] for k in order])
    vec = vec.subs(mi[4], m[1] * m[2])
    vec = vec.subs(mi[5], m[1] ** 2 - m[2] ** 2)
    vec = vec.subs(mi[6], 2 * m[4] * m[1] + (m[3] + THIRD) * m[2] - 2 * m[2] * m[1] ** 2)
    vec = vec.subs(mi[7], 2 * m[4] * m[2] + (m[5] + THIRD) * m[1] - 2 * m[1] * m[2] ** 2)
    vec = vec.subs(m[0], 1)
    fn = sp.lambdify([m[k] for k in range(1, 9)], list(vec), 'numpy')
    return fn, [str(k) for k in order]
