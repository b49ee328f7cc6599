"""Host-side surface of the drop-in classes without a GPU: RootLB's sizes,
strides, V and their dtypes (the checks of the reference's own
test_RootLBVars, idpy/LBM/test.py:318-362), and the lexicographic index
helpers IndexFromPos / PosFromIndex (idpy/LBM/MultiPhase.py:94-106)."""
import numpy as np
import pytest


@pytest.mark.parametrize('dims,xi', [((11, 7), 'D2Q9'), ((5, 6, 7), 'D3Q19')])
def test_rootlb_sizes_strides_and_volume(dims, xi):
    from idmx.LBM.LBM import LBMTypes, RootLB, XIStencils
    r = RootLB(dim_sizes=dims, xi_stencil=XIStencils[xi], custom_types=LBMTypes)
    v = r.sims_vars
    assert list(v['dim_sizes']) == list(dims)
    assert v['dim_sizes'].dtype == np.int32 and v['dim_strides'].dtype == np.int32
    strides = [int(np.prod(dims[:k + 1])) for k in range(len(dims) - 1)]
    assert list(v['dim_strides']) == strides
    assert v['V'] == int(np.prod(dims))
    assert v['Q'] == (9 if xi == 'D2Q9' else 19) and v['DIM'] == len(dims)


@pytest.mark.parametrize('dims', [(11, 7), (5, 6, 7)])
def test_index_position_round_trip(dims):
    """Every site: PosFromIndex inverts IndexFromPos, and the index is
    x + L0 (y + L1 z), x fastest."""
    from idmx.LBM.MultiPhase import IndexFromPos, PosFromIndex
    strides = [int(np.prod(dims[:k + 1])) for k in range(len(dims) - 1)]
    V = int(np.prod(dims))
    for g in range(V):
        pos = PosFromIndex(g, strides)
        assert all(0 <= p < L for p, L in zip(pos, dims))
        lex = pos[0] + dims[0] * (pos[1] + (dims[1] * pos[2] if len(dims) == 3 else 0))
        assert lex == g
        assert IndexFromPos(pos, strides) == g


def test_relaxation_matrix_moments_are_the_bases():
    """RelaxationMatrix (idpy/LBM/Collisions.py:41-84) on the drop-in: its
    moment polynomials evaluated on the velocity set are the rows of the
    W-orthogonal basis (default) and of the non-orthogonal invertible basis
    (WOrth_flag=False), exactly; GetIndexFromMoment / SetOmegasByMoment /
    GetDiagonalMatrix index rows in that order."""
    import sympy as sp
    from fractions import Fraction
    from idmx.LBM import hermite
    from idmx.LBM.Collisions import RelaxationMatrix
    from idmx.LBM.Equilibria import HermiteEquilibria
    from idmx.LBM.stencils import XIStencils
    for name in ('D2Q9', 'D3Q19'):
        eq = HermiteEquilibria(xi_stencil=XIStencils[name])
        xis = XIStencils[name]['XIs']
        for worth, rows in ((True, hermite.worth_set(name)[0]),
                            (False, hermite.invertible_set(name)[0])):
            rm = RelaxationMatrix(eq, WOrth_flag=worth)
            ms = rm.GetMoments()
            assert len(ms) == eq.Q
            syms = sorted(set().union(*[m.free_symbols for m in ms]), key=str)
            for k, m in enumerate(ms):
                vals = [m.subs(dict(zip(syms, v))) for v in xis] if syms else [m] * len(xis)
                assert [Fraction(int(sp.Rational(v).p), int(sp.Rational(v).q)) for v in vals] \
                    == list(rows[k]), (name, worth, k)
                assert rm.GetIndexFromMoment(m) == k
        rm = RelaxationMatrix(eq)
        om = sp.symbols('omega_0:%d' % eq.Q)
        S = rm.GetDiagonalMatrix(list(om))
        assert S.shape == (eq.Q, eq.Q) and S[3, 3] == om[3] and S[3, 4] == 0
        ms = rm.GetMoments()
        S2 = rm.SetOmegasByMoment({(ms[4], ms[4]): 1.2, (ms[4], ms[5]): 0.1})
        assert S2[4, 4] == 1.2 and S2[4, 5] == 0.1 and sum(abs(x) for x in S2) == 1.3
