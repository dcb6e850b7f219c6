"""MI355X-native (gfx950) finite-volume update of the Cosmos hydrodynamics code.

The compute path is the HIP library libcosmos_fv.so (C-ABI: include/cosmos_fv.h);
this package is its Python binding, the build script and the seeded workloads.
"""
from .native import CosmosError, LIB_PATH, load  # noqa: F401
from .batch import Batch  # noqa: F401
from .solver import (BLACKHOLE, EXACT, FAST, MINBEE, NEWTONIAN, RELATIVISTIC, SUPERBEE, TAPE_NONE,  # noqa: F401
                     TAPE_RERUN, TAPE_SEGMENT, VANLEER, Adjoint, Amr, HostArray, Solver, host_empty)

__all__ = ["Solver", "HostArray", "host_empty", "Amr", "Adjoint", "Batch", "CosmosError", "BLACKHOLE", "load", "LIB_PATH", "NEWTONIAN", "RELATIVISTIC", "SUPERBEE", "VANLEER",
           "MINBEE"]
