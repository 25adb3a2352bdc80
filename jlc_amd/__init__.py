"""jlc_amd: MI355X-native DGEMM engine for the JLC linear-algebra library.

The product is libjlc_hip.so (C ABI in include/jlc_hip.h, HIP kernels in
jlc_amd/csrc/). This package holds its ctypes binding (native) and the
host-side mirror of the reference's GEMM API (compute).
"""
from . import native  # noqa: F401

__all__ = ["native"]
