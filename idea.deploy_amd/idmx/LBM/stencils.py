"""Velocity sets and isotropic-gradient E-stencils.

Same tables, names and layout as idpy's XIStencils / FStencils
(idpy/LBM/LBM.py:74-185): dicts with 'XIs'/'Es' vector tuples, 'Ws' weights
as Python floats and 'c2'. Notebooks index them exactly as before, e.g.
XIStencils['D2Q9'], FStencils['D2E8'].
"""
from collections import defaultdict

XIStencils = defaultdict(lambda: defaultdict(dict))
FStencils = defaultdict(lambda: defaultdict(dict))

XIStencils['D2Q9'] = {
    'XIs': ((0, 0), (1, 0), (0, 1), (-1, 0), (0, -1),
            (1, 1), (-1, 1), (-1, -1), (1, -1)),
    'Ws': (4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36),
    'c2': 1 / 3}

XIStencils['D3Q19'] = {
    'XIs': ((0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0),
            (1, 0, 1), (-1, 0, 1), (-1, 0, -1), (1, 0, -1),
            (0, 1, 1), (0, -1, 1), (0, -1, -1), (0, 1, -1)),
    'Ws': (1 / 3,) + (1 / 18,) * 6 + (1 / 36,) * 12,
    'c2': 1 / 3}


# Order of vectors within each shell follows idpy/LBM/LBM.py:94-185.
FStencils['D2E4'] = {
    'Es': ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)),
    'Ws': (1 / 3,) * 4 + (1 / 12,) * 4}

FStencils['D2E6'] = {
    'Es': ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1),
           (2, 0), (0, 2), (-2, 0), (0, -2)),
    'Ws': (4 / 15,) * 4 + (1 / 10,) * 4 + (1 / 120,) * 4}

FStencils['D2E8'] = {
    'Es': ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1),
           (2, 0), (0, 2), (-2, 0), (0, -2),
           (2, 1), (1, 2), (-1, 2), (-2, 1), (-2, -1), (-1, -2), (1, -2), (2, -1),
           (2, 2), (-2, 2), (-2, -2), (2, -2)),
    'Ws': (4 / 21,) * 4 + (4 / 45,) * 4 + (1 / 60,) * 4 + (2 / 315,) * 8 + (1 / 5040,) * 4}

_D3_SHELL1 = ((1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
_D3_SHELL2 = ((1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0),
              (0, 1, 1), (0, -1, 1), (0, -1, -1), (0, 1, -1),
              (1, 0, 1), (1, 0, -1), (-1, 0, -1), (-1, 0, 1))
_D3_SHELL3 = ((1, 1, 1), (-1, 1, 1), (-1, -1, 1), (1, -1, 1),
              (1, 1, -1), (-1, 1, -1), (-1, -1, -1), (1, -1, -1))
_D3_SHELL4 = ((2, 0, 0), (0, 2, 0), (-2, 0, 0), (0, -2, 0), (0, 0, 2), (0, 0, -2))
_D3_SHELL5 = ((2, 1, 0), (2, -1, 0), (2, 0, 1), (2, 0, -1),
              (1, 2, 0), (-1, 2, 0), (0, 2, 1), (0, 2, -1),
              (-2, 1, 0), (-2, -1, 0), (-2, 0, 1), (-2, 0, -1),
              (1, -2, 0), (-1, -2, 0), (0, -2, 1), (0, -2, -1),
              (1, 0, 2), (-1, 0, 2), (0, 1, 2), (0, -1, 2),
              (1, 0, -2), (-1, 0, -2), (0, 1, -2), (0, -1, -2))
_D3_SHELL6 = ((2, 1, 1), (2, -1, 1), (2, -1, -1), (2, 1, -1),
              (1, 2, 1), (-1, 2, 1), (-1, 2, -1), (1, 2, -1),
              (-2, 1, 1), (-2, -1, 1), (-2, -1, -1), (-2, 1, -1),
              (1, -2, 1), (-1, -2, 1), (-1, -2, -1), (1, -2, -1),
              (1, 1, 2), (-1, 1, 2), (-1, -1, 2), (1, -1, 2),
              (1, 1, -2), (-1, 1, -2), (-1, -1, -2), (1, -1, -2))
_D3_SHELL8 = ((2, 2, 0), (-2, 2, 0), (-2, -2, 0), (2, -2, 0),
              (2, 0, 2), (-2, 0, 2), (-2, 0, -2), (2, 0, -2),
              (0, 2, 2), (0, -2, 2), (0, -2, -2), (0, 2, -2))

FStencils['D3E4'] = {
    'Es': _D3_SHELL1 + _D3_SHELL2,
    'Ws': (1 / 6,) * 6 + (1 / 12,) * 12}

FStencils['D3E6'] = {
    'Es': _D3_SHELL1 + _D3_SHELL2 + _D3_SHELL3 + _D3_SHELL4,
    'Ws': (2 / 15,) * 6 + (1 / 15,) * 12 + (1 / 60,) * 8 + (1 / 120,) * 6}

FStencils['D3E8'] = {
    'Es': _D3_SHELL1 + _D3_SHELL2 + _D3_SHELL3 + _D3_SHELL4 + _D3_SHELL5 + _D3_SHELL6 + _D3_SHELL8,
    'Ws': ((4 / 45,) * 6 + (1 / 21,) * 12 + (2 / 105,) * 8 + (5 / 504,) * 6
           + (1 / 315,) * 24 + (1 / 630,) * 24 + (1 / 5040,) * 12)}


def estencil_name(f_stencil):
    """Name ('D2E4', ...) of a stencil dict whose vectors and weights equal one
    of FStencils (weights compared to 1e-12, so sympy Rationals match too)."""
    es = tuple(tuple(int(c) for c in e) for e in f_stencil.get('Es', f_stencil.get('XIs')))
    ws = tuple(float(w) for w in f_stencil['Ws'])
    for name, st in FStencils.items():
        ref = dict(zip(st['Es'], st['Ws']))
        if len(ref) == len(es) and all(e in ref and abs(ref[e] - w) < 1e-12 for e, w in zip(es, ws)):
            return name
    return None


def xistencil_name(xi_stencil):
    xis = tuple(tuple(int(c) for c in v) for v in xi_stencil['XIs'])
    ws = tuple(float(w) for w in xi_stencil['Ws'])
    for name, st in XIStencils.items():
        if st['XIs'] == xis and all(abs(a - b) < 1e-12 for a, b in zip(st['Ws'], ws)):
            return name
    return None
