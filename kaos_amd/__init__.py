"""kaos_amd -- MI355X (gfx950) implementation of the KAOS per-timestep kinetic
hot path behind a C-ABI (include/kaos_hip.h, libkaos_hip.so).

Modules:
  _abi    ctypes declarations of the C-ABI and the library loader
  solver  KineticSolver: host driver that mirrors KineticSolverSub's calls
  grids   flux-tube grids and initial ions of the benchmark/test workloads
  build   hipcc build of libkaos_hip.so (and the Fortran interface/driver)
"""
from ._abi import KAOS_FP64, KAOS_MIXED, load  # noqa: F401
