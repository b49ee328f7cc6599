"""Lattice Boltzmann simulation classes (idpy.LBM surface)."""
