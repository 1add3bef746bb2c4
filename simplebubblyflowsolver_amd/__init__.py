"""MI355X-native (gfx950 HIP) two-phase time step of SimpleBubblyFlowSolver.

The product is libsbfs.so: a C-ABI library exporting the reference's entry
points (integrate, decide_dt, fluid_*, interface_*), declared in include/.
This package holds its HIP sources (csrc/), the build script and a thin
ctypes host mirror (solver.py) used by tests and the benchmark.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
