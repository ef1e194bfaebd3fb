"""enlightning_amd -- MI355X-native per-patch update path of Enlightning.

The product is libenlightning_hip.so (HIP kernels for gfx950 behind the C ABI
of include/enlightning_hip.h).  This package builds it and exposes thin
Python calls for tests and benchmarks.
"""
from ._lib import VARIANTS, VARIANT_ID, load, LIB_PATH  # noqa: F401
from .gas import Gas, RK3  # noqa: F401
from .workloads import Workload, SMALL_256, PULSES_4096, initial_fields  # noqa: F401

__all__ = ["VARIANTS", "VARIANT_ID", "load", "LIB_PATH", "Gas", "RK3", "Workload", "SMALL_256",
           "PULSES_4096", "initial_fields"]
