"""MI355X-native CTF estimation (GCtfFind's CTF-estimation path).

libgctf_hip.so (csrc/) holds the HIP kernels and the C-ABI declared in
include/gctf_hip.h; gctf.py binds it with ctypes.
"""
from .gctf import (CFindCtf2D, Context, GctfError, load, LIB_PATH, DTYPE_F16, DTYPE_F32,  # noqa: F401
                   EXPORTS)
