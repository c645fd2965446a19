"""vflip_amd: MI355X-native (gfx950) vFLIP per-step hot path.

The compute path is libvflip.so (hand-written HIP kernels behind the C ABI in
include/vflip.h). Importing the package loads it and raises when the HIP build
is missing; there is no CPU fallback.
"""
from ._lib import LIB_PATH, VflipError, check, lib  # noqa: F401

__all__ = ["lib", "check", "VflipError", "LIB_PATH"]
