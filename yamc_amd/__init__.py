"""MI355X-native Ewald Monte Carlo energy/force path of YAMC.

The compute lives in libyamc_ewald.so (HIP kernels for gfx950 behind the C ABI
of include/yamc_ewald.h).  Importing this package does not touch the GPU; the
library is loaded on first use and its absence is an error, never a fallback.
"""
from ._lib import YAMC_FP32C, YAMC_FP64, YamcError, YamcLibraryError, load  # noqa: F401
from .system import CL_DANG, NA_DANG, SPCE, Params, System, build_system, config, replica_system  # noqa: F401

__all__ = ["YAMC_FP32C", "YAMC_FP64", "YamcError", "YamcLibraryError", "load", "Params", "System", "build_system",
           "config", "replica_system", "SPCE", "NA_DANG", "CL_DANG"]
