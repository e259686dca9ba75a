"""mag2d_amd: MI355X-native PIC-MCC inner loop of mag2d behind a C ABI.

The compute path is libmag2d_amd.so (hand-written HIP kernels for gfx950,
include/mag2d_amd.h); ``mag2d_amd.abi`` is its ctypes binding and
``mag2d_amd/host/mag2d.hpp`` the C++ class layer that mirrors the
reference's Species / Field2D / Fields / t_grid API.
"""
__all__ = ["abi"]
