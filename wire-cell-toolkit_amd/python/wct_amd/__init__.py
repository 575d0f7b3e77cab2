"""Python bindings (ctypes) of the gfx950 HIP library's C ABI
(include/wct_hip.h).  Device buffers are torch tensors on cuda:0;
torch is only the allocator and stream provider here.

The library must be built (wire-cell-toolkit_amd/libwct_hip.so); there is
no fallback: loading fails loudly when it is missing.
"""
from .capi import lib, check, WctError, LIB_PATH  # noqa: F401
