"""yallasql_amd — the YallaSQL operator hot path on AMD MI355X (gfx950).

The product is the C ABI in include/ysql_kernels.h, implemented by the HIP
kernels of libyallasql_amd.so.  `yallasql_amd.ops` exposes it on torch
tensors; `yallasql_amd._lib` is the raw ctypes binding.
"""
import os

__all__ = ["lib_path", "version"]


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libyallasql_amd.so")


def version():
    from ._lib import lib

    return lib.ysql_version().decode()
