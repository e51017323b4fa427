"""MI355X-native moment-representation lattice Boltzmann solver (D3Q19 / D3Q27).

The compute path is the HIP library liblbm_amd.so (C ABI: include/lbm_amd.h);
this package holds its sources (csrc/), the build recipe and the host-side
binding that mirrors the reference's case interface.
"""
from .native import Batch, LBMError, Solver, Surrogate, autotune  # noqa: F401

__all__ = ["Solver", "Batch", "Surrogate", "autotune", "LBMError"]
