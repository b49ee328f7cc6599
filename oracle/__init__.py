"""CPU oracle for the Shan-Chen LBM step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import anything here, and only to check results (or as the timed CPU
baseline), never as the product path. The product library (idea.deploy_amd/idmx,
libidmx.so) does not import, link or execute this package.

Modules
  stencils     velocity sets and E-stencils as in idpy/LBM/LBM.py:74-185
  cpuref       ctypes wrapper of libcpuref.so (cpu_ref.cpp), the serial C++
               restatement of the reference's CTypes kernels
  np_ref       numpy brute force (np.roll) of one step, for tiny grids
  hermite_mrt  sympy derivation of the MRT bases (exact rationals)
  iso_filter   sympy derivation of the D2Q9 isotropy filter's tables and
               the reference's symbolic solve of it
  nn_ref       numpy forward / backward of the warm-start network

Parity status of each function is listed in the module docstrings; the pins
live in tests/test_oracle_*.py.
"""
