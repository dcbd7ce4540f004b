"""MI355X-native engine for the checkerboard Metropolis sweep of McLics-style nematic lattices.

The compute path is libmclgpu (hand-written HIP kernels for gfx950 behind the C ABI of include/mclgpu.h); this
package holds its sources (csrc/), the host program mclics (host/), the ctypes binding (capi) and the build
script (build).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
