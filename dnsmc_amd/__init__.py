"""MI355X-native Monte Carlo radiative-transfer solver for the channel DNS.

The product is libdnsmc_rad.so (HIP kernels for gfx950 + the C ABI of
include/dnsmc_rad.h).  This package is its host-side mirror: a ctypes
binding (``Solver``, ``mc_gpu``, ``get_results``), the build script and the
synthetic workload generators.
"""
from .solver import (Solver, config_from_dict, default_config, fix_rad_gpu, fix_temp_gpu, get_results,  # noqa: F401
                     mc_gpu, mc_gpu_commit, parse_config, probe_tables, sync)
from ._lib import DnsmcError  # noqa: F401

__all__ = ["Solver", "config_from_dict", "default_config", "get_results", "mc_gpu", "parse_config", "sync",
           "fix_temp_gpu", "mc_gpu_commit", "fix_rad_gpu", "probe_tables", "DnsmcError"]
