"""MI355X (gfx950) implementation of the RUMDEED per-step particle hot path.

The compute lives in librumdeed.so (HIP kernels behind a C ABI,
include/rumdeed.h); this package is the Python host side over that ABI.
"""
__all__ = ["diode", "workloads"]
