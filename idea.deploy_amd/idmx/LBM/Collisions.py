"""idpy.LBM.Collisions surface: RelaxationMatrix (idpy/LBM/Collisions.py:
41-84), the helper notebooks use to build the relaxation matrix S that
MainLoopMRT takes. The moments are the package's exact Hermite basis
(idmx/LBM/hermite.py): the W-orthogonal Gram-Schmidt set by default, the
non-orthogonal invertible set with WOrth_flag=False, as sympy polynomials in
the symbols \\xi_0 .. \\xi_{D-1}."""
import numpy as np
import sympy as sp

from . import hermite


class RelaxationMatrix:
    def __init__(self, eq_obj=None, search_depth=6, WOrth_flag=True):
        if eq_obj is None:
            raise Exception("Missing argument 'eq_obj'")
        self.D, self.Q = eq_obj.D, eq_obj.Q
        self.MHermitePolys = np.empty(self.Q, dtype=object)
        polys = hermite.moment_polys(eq_obj.xi_name, worth=bool(WOrth_flag))
        if len(polys) != self.Q:
            raise Exception('moment basis has %d rows for Q = %d' % (len(polys), self.Q))
        for i, p in enumerate(polys):
            self.MHermitePolys[i] = p

    def GetMoments(self):
        return self.MHermitePolys

    def DisplayMoments(self):
        """Prints each moment as LaTeX (the reference displays them through
        IPython when it is available)."""
        try:
            from IPython.display import Latex, display
        except ImportError:
            display = None
        for i, m in enumerate(self.MHermitePolys):
            text = str(i) + ' :$' + sp.latex(m) + '$'
            if display is not None:
                display(Latex(text))
            else:
                print(text)

    def GetIndexFromMoment(self, moment):
        for i, m in enumerate(self.MHermitePolys):
            if sp.expand(m - moment) == 0:
                return i
        raise IndexError('moment %s is not in the basis' % (moment,))

    def GetDiagonalMatrix(self, omega_list=None):
        self.omega_list = omega_list
        return sp.Matrix.diag(omega_list)

    def SetOmegasByMoment(self, omega_dict=None):
        S = sp.zeros(self.Q)
        for key, val in omega_dict.items():
            S[self.GetIndexFromMoment(key[0]), self.GetIndexFromMoment(key[1])] = val
        return S
