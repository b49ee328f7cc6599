"""The psi_code compiler: C semantics of the expressions it accepts (integer
folding, float rounding, casts, precedence), checked against the same
expression compiled by gcc and evaluated by an interpreter that follows the
device's stack machine op for op; and the class surface routing a psi_code
outside the closed forms to it."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
import sympy as sp

from idmx.LBM import psi_code as pc

EXPRS = ['1. - exp(-ln_0)', 'exp(-1./ln_0)', 'ln_0', '1/2*ln_0 + 1. - exp(-ln_0)',
         '(1/2.)*ln_0', 'pow(ln_0, 1.5) - 0.25f*ln_0', 'sqrt(fabs(ln_0 - 1)) * (float)ln_0',
         '-ln_0*ln_0/3 + 2*ln_0', 'tanh(2*(ln_0 - 0.5))', 'fmax(ln_0, 0.3) - fmin(ln_0, 0.7)',
         '(double)ln_0 * (NType)2', '3 - 7/2 * ln_0', '-(-3)/2 + ln_0/-4',
         'log(1 + ln_0) * cos(ln_0) + sin(ln_0) * sinh(0.1*ln_0) / cosh(ln_0)',
         '2.5e-1*ln_0 - 1e1/ln_0', '(PsiType)(1. - exp(-ln_0))*(float)1.1']


def _interpret(ops, consts, n):
    """The device's stack machine (csrc/sc_kernels.h psi_prog), op for op."""
    f32 = lambda x: float(np.float32(x))
    un = {pc.NEG: lambda b: -b, 7: math.exp, 8: math.log, 9: math.sqrt, 11: math.fabs,
          12: math.tanh, 13: math.sinh, 14: math.cosh, 15: math.sin, 16: math.cos, pc.F32: f32}
    bi = {pc.ADD: lambda a, b: a + b, pc.SUB: lambda a, b: a - b, pc.MUL: lambda a, b: a * b,
          pc.DIV: lambda a, b: a / b, 10: math.pow, 17: max, 18: min}
    st, ci = [], 0
    for op in ops:
        if op == pc.N:
            st.append(n)
        elif op == pc.C:
            st.append(float(consts[ci]))
            ci += 1
        elif op in un:
            st[-1] = un[op](st[-1])
        else:
            b = st.pop()
            st[-1] = bi[op](st[-1], b)
    assert len(st) == 1 and ci == len(consts)
    return st[0]


@pytest.fixture(scope='module')
def gcc_funcs(tmp_path_factory):
    d = tmp_path_factory.mktemp('psi')
    src = d / 'p.c'
    body = ['#include <math.h>', 'typedef double NType; typedef double PsiType;']
    for i, e in enumerate(EXPRS):
        body.append('double psi%d(double ln_0) { return %s; }' % (i, e))
    src.write_text('\n'.join(body) + '\n')
    so = d / 'p.so'
    subprocess.run(['gcc', '-std=c99', '-O3', '-ffp-contract=off', '-fPIC', '-shared', '-o',
                    str(so), str(src), '-lm'], check=True)
    L = C.CDLL(str(so))
    out = []
    for i in range(len(EXPRS)):
        f = getattr(L, 'psi%d' % i)
        f.restype, f.argtypes = C.c_double, [C.c_double]
        out.append(f)
    return L, out


@pytest.mark.parametrize('i', range(len(EXPRS)))
def test_program_equals_gcc(gcc_funcs, i):
    _, funcs = gcc_funcs
    ops, consts = pc.compile_psi(EXPRS[i])
    for n in np.linspace(0.05, 2.5, 97):
        want = funcs[i](float(n))
        got = _interpret(ops, consts, float(n))
        assert got == want or (math.isnan(got) and math.isnan(want)), (EXPRS[i], n, got, want)


def test_integer_semantics_are_folded():
    ops, consts = pc.compile_psi('1/2*ln_0')
    assert _interpret(ops, consts, 1.7) == 0.0
    ops, consts = pc.compile_psi('3 - 7/2 * ln_0')
    assert list(consts) == [3.0, 3.0]
    ops, consts = pc.compile_psi('-7/2')
    assert list(ops) == [pc.C] and list(consts) == [-3.0]  # C truncates toward zero


@pytest.mark.parametrize('bad', ['exp(ln_0', 'foo(ln_0)', 'ln_0 +', 'x * 2', 'pow(ln_0)',
                                 'ln_0 ? 1 : 0', ' + '.join(['ln_0'] * 40)])
def test_rejects_outside_the_grammar(bad):
    with pytest.raises(NotImplementedError):
        pc.compile_psi(bad)


def test_class_routes_other_psi_code_to_a_program():
    from idmx.LBM._base import psi_kind_from
    n = sp.Symbol('n')
    assert psi_kind_from(1 - sp.exp(-n), '1. - exp(-ln_0)') == 'exp'
    assert psi_kind_from(sp.sqrt(n), 'sqrt(ln_0)') == 'program'
    with pytest.raises(NotImplementedError):
        psi_kind_from(sp.sqrt(n), 'sqrt(ln_0) ? 1 : 0')
