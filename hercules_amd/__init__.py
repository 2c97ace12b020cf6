"""MI355X exact k-NN query engine for Hercules data-series indexes.

The compute path is libhkg.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/hkg.h); this package is its Python binding.
"""
from .hkg import (HerculesIndex, Server, DeviceArray, HkgError, KnnResult, breakpoints256, gen_series,
                  gen_noisy_queries, gen_series_device, lib, FLT_MAX, NO_ID)

__all__ = ["HerculesIndex", "Server", "DeviceArray", "HkgError", "KnnResult", "breakpoints256", "gen_series",
           "gen_noisy_queries", "gen_series_device", "lib", "FLT_MAX", "NO_ID"]
