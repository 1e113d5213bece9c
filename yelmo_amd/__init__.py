"""yelmo_amd -- MI355X-native SSA/DIVA velocity solve of the Yelmo ice-sheet model.

The solver itself is libygpu.so (HIP, gfx950), reached through the C-ABI of
include/ygpu.h; this package holds its ctypes binding, the host-side mirror of
the reference's velocity-solver interface, the input-side preparation and the
batched / served drivers.
"""
from .params import SsaParams, benchmark_params  # noqa: F401

__all__ = ["SsaParams", "benchmark_params"]
