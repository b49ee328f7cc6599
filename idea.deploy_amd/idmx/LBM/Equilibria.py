"""idpy.LBM.Equilibria surface: HermiteEquilibria as the stencil carrier the
MRT helpers take (idpy/LBM/Equilibria.py:53-138). The equilibrium itself is
evaluated by the HIP kernels (second-order Hermite, exact constants)."""
from .stencils import xistencil_name


class HermiteEquilibria:
    """Carries the velocity set of an MRT collision: D, Q and the stencil
    (RelaxationMatrix reads them from here, as on the reference)."""

    def __init__(self, xi_stencil=None, root_pop='f', root_n='n', root_u='u', root_w='w'):
        if xi_stencil is None:
            raise Exception("Missing argument 'xi_stencil'")
        self.xi_stencil = xi_stencil
        self.xi_name = xistencil_name(xi_stencil)
        self.D = len(xi_stencil['XIs'][0])
        self.Q = len(xi_stencil['XIs'])
        self.root_pop, self.root_n, self.root_u, self.root_w = root_pop, root_n, root_u, root_w
