"""idmx -- MI355X-native Shan-Chen lattice Boltzmann behind idpy's surface.

    from idmx.IdpyCode import HIP_T
    from idmx.LBM.MultiPhase import ShanChenMultiPhase   # Meta class
    from idmx.LBM.LBM import ShanChenMultiPhase          # legacy class

All compute runs in libidmx.so (hand-written HIP for gfx950) through the C
ABI of include/idmx.h; the Python layer only builds parameters and moves
device pointers.
"""
from ._lib import IdmxError, lib  # noqa: F401

__version__ = '0.1.0'
