"""tfg_amd - MI355X (gfx950) implementation of the TFG path tracer hot path.

The compute path is libtfg_hip.so (hand-written HIP kernels behind the C ABI
of include/tfg.h); this package holds its ctypes binding, the seeded scene
and ray-batch generators, and the host-side frame routines.
"""
from ._lib import TfgError, lib  # noqa: F401

__all__ = ["TfgError", "lib"]
