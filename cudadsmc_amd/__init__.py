"""MI355X-native DSMC step library (gfx950 HIP kernels behind a C ABI).

The product is cudadsmc_amd/libdsmc.so (include/dsmc.h).  This package holds
its build recipe, a ctypes view of the ABI and `Simulation`, the host-side
mirror of the reference driver's call sequence.
"""
from . import _abi, configs
from ._abi import DsmcError, load_library
from .configs import Run, Trap
from .simulation import OBS_DTYPE, Simulation, SimulationGroup

__all__ = ["DsmcError", "load_library", "Run", "Trap", "Simulation", "SimulationGroup", "OBS_DTYPE",
           "configs", "_abi"]
