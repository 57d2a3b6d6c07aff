"""MI355X pressure-Poisson multigrid for the ghost-cell IBM flow solver.

The product is the C-ABI library libpp_hip.so (include/pp_hip.h); this
package holds its sources (csrc/), the Fortran binding (fortran/) and the
ctypes mirror used from Python (pp.py).
"""
from . import pp  # noqa: F401
