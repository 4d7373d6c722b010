"""MI355X (gfx950) implementation of RTGL1's ray-traversal and denoising path.

The product is the C-ABI library librtgl1_amd.so (include/RTGL1/RTGL1.h and
RTGL1_ext.h); this package holds its build script, a ctypes binding and the
seeded synthetic scenes of the benchmark workloads.
"""
from .ffi import lib, device_count  # noqa: F401

__all__ = ["lib", "device_count"]
