"""Compiles a pseudopotential given as a C expression (the psi_code kwarg)
into the stack program of idmx_sc_set_psi_program.

The reference pastes psi_code into the body of K_ComputeDensityPsiMeta
(idpy/LBM/MultiPhaseKernelsMeta.py:107-167; idpy/LBM/MultiPhase.py:2120)
and lets the C compiler evaluate it. This module parses the same subset of
C expressions and keeps C's arithmetic: integer literals and integer
sub-expressions are folded with C's integer rules at compile time (1/2 is
0), an operation with a double operand is done in double, an f-suffixed
literal or a (float) cast rounds to single precision; (double), (NType),
(PsiType), (UType) and (PopType) casts are no-ops.

Grammar: expression := term (('+' | '-') term)*; term := unary (('*' | '/')
unary)*; unary := ('-' | '+' | cast) unary | primary; primary := number |
variable | function '(' args ')' | '(' expression ')'. Functions: exp, log,
sqrt, pow, fabs, tanh, sinh, cosh, sin, cos, fmax, fmin.
"""
import re

import numpy as np

# op codes (csrc/sc_kernels.h PsiOp)
N, C, ADD, SUB, MUL, DIV, NEG = 0, 1, 2, 3, 4, 5, 6
FUNCS1 = {'exp': 7, 'log': 8, 'sqrt': 9, 'fabs': 11, 'tanh': 12, 'sinh': 13, 'cosh': 14,
          'sin': 15, 'cos': 16}
FUNCS2 = {'pow': 10, 'fmax': 17, 'fmin': 18}
F32 = 19
MAX_OPS, MAX_CONSTS, MAX_STACK = 48, 16, 8
_NOOP_CASTS = ('double', 'NType', 'PsiType', 'UType', 'PopType')

_TOKEN = re.compile(r'\s*(?:(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[fF]?)'
                    r'|(?P<id>[A-Za-z_]\w*)|(?P<op>[-+*/(),]))')


def _tokens(text):
    pos, out = 0, []
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise NotImplementedError('psi_code: cannot parse %r at %r' % (text, text[pos:]))
        pos = m.end()
        out.append((m.lastgroup, m.group(m.lastgroup)))
    return out


class _Node:
    """An expression node: kind 'int' (a folded C int), 'dbl' (a double
    constant), or 'code' (ops that leave one double on the stack)."""

    def __init__(self, kind, value=None, ops=None, consts=None, depth=1):
        self.kind, self.value = kind, value
        self.ops, self.consts, self.depth = ops or [], consts or [], depth

    def code(self):
        if self.kind == 'code':
            return self.ops, self.consts, self.depth
        return [C], [float(self.value)], 1


def _c_int_div(a, b):
    if b == 0:
        raise NotImplementedError('psi_code: integer division by zero')
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _binary(op, a, b):
    if a.kind == 'int' and b.kind == 'int':
        x, y = a.value, b.value
        v = {'+': x + y, '-': x - y, '*': x * y}.get(op)
        return _Node('int', _c_int_div(x, y) if op == '/' else v)
    if a.kind != 'code' and b.kind != 'code':
        x, y = float(a.value), float(b.value)
        v = {'+': x + y, '-': x - y, '*': x * y, '/': x / y if y != 0 else None}[op]
        if v is not None:
            return _Node('dbl', v)
    ao, ac, ad = a.code()
    bo, bc, bd = b.code()
    code = {'+': ADD, '-': SUB, '*': MUL, '/': DIV}[op]
    return _Node('code', ops=ao + bo + [code], consts=ac + bc, depth=max(ad, bd + 1))


class _Parser:
    def __init__(self, toks, names):
        self.t, self.i, self.names = toks, 0, names

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if tok[0] is None or (val is not None and tok[1] != val):
            raise NotImplementedError('psi_code: expected %r, found %r' % (val, tok[1]))
        self.i += 1
        return tok

    def expression(self):
        node = self.term()
        while self.peek()[1] in ('+', '-'):
            op = self.take()[1]
            node = _binary(op, node, self.term())
        return node

    def term(self):
        node = self.unary()
        while self.peek()[1] in ('*', '/'):
            op = self.take()[1]
            node = _binary(op, node, self.unary())
        return node

    def unary(self):
        kind, val = self.peek()
        if val == '-':
            self.take()
            x = self.unary()
            if x.kind != 'code':
                return _Node(x.kind, -x.value)
            return _Node('code', ops=x.ops + [NEG], consts=x.consts, depth=x.depth)
        if val == '+':
            self.take()
            return self.unary()
        if val == '(' and self.i + 2 < len(self.t) and self.t[self.i + 1][0] == 'id' \
                and self.t[self.i + 2][1] == ')' \
                and self.t[self.i + 1][1] in _NOOP_CASTS + ('float', 'int'):
            cast = self.t[self.i + 1][1]
            self.i += 3
            x = self.unary()
            if cast in _NOOP_CASTS:
                return x if x.kind != 'int' else _Node('dbl', float(x.value))
            if cast == 'int':
                if x.kind == 'code':
                    raise NotImplementedError('psi_code: (int) of a density-dependent value')
                return _Node('int', int(x.value))
            if x.kind != 'code':
                return _Node('dbl', float(np.float32(x.value)))
            return _Node('code', ops=x.ops + [F32], consts=x.consts, depth=x.depth)
        return self.primary()

    def primary(self):
        kind, val = self.take()
        if kind == 'num':
            if val[-1] in 'fF':
                return _Node('dbl', float(np.float32(float(val[:-1]))))
            if re.fullmatch(r'\d+', val):
                return _Node('int', int(val))
            return _Node('dbl', float(val))
        if kind == 'id':
            if val in self.names:
                return _Node('code', ops=[N], depth=1)
            if val in FUNCS1 or val in FUNCS2:
                self.take('(')
                args = [self.expression()]
                while self.peek()[1] == ',':
                    self.take(',')
                    args.append(self.expression())
                self.take(')')
                nargs = 1 if val in FUNCS1 else 2
                if len(args) != nargs:
                    raise NotImplementedError('psi_code: %s takes %d arguments' % (val, nargs))
                ops, consts, depth = [], [], 0
                for k, a in enumerate(args):
                    ao, ac, ad = a.code()
                    ops, consts, depth = ops + ao, consts + ac, max(depth, ad + k)
                ops.append(FUNCS1.get(val, FUNCS2.get(val)))
                return _Node('code', ops=ops, consts=consts, depth=depth)
            raise NotImplementedError('psi_code: unknown name %r' % val)
        if val == '(':
            node = self.expression()
            self.take(')')
            return node
        raise NotImplementedError('psi_code: unexpected %r' % val)


def compile_psi(psi_code, names=('ln_0', 'ln', 'n')):
    """(ops uint8 array, consts float64 array) of psi_code. Raises
    NotImplementedError for what the grammar above does not cover."""
    p = _Parser(_tokens(psi_code), names)
    node = p.expression()
    if p.i != len(p.t):
        raise NotImplementedError('psi_code: trailing %r' % (p.t[p.i][1],))
    ops, consts, depth = node.code()
    if len(ops) > MAX_OPS or len(consts) > MAX_CONSTS or depth > MAX_STACK:
        raise NotImplementedError('psi_code: program too long (%d ops, %d constants, depth %d)'
                                  % (len(ops), len(consts), depth))
    return np.array(ops, dtype=np.uint8), np.array(consts, dtype=np.float64)
