"""afepack_amd: MI355X (gfx950) assemble-and-solve path of AFEPack.

The product is the C ABI in include/afepack_amd.h, implemented by the in-tree
shared library libafepack_amd.so (HIP kernels + host C++).  This package only
loads that library with ctypes and wraps it for Python callers; device vectors
are DeviceVector buffers allocated through the library itself.

Importing does not need a GPU.  Every compute entry point goes through the
library; when the library is missing, load() raises -- there is no fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libafepack_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "afepack_amd.h")

_lib = None


class AfeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"afepack_amd error {code}: {msg}")
        self.code = code


def load():
    """Load libafepack_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python afepack_amd/build.py` "
            "(the HIP path is the only implementation)")
    lib = ctypes.CDLL(LIB_PATH)
    from . import _sigs
    _sigs.declare(lib)
    _lib = lib
    return lib


def check(code):
    if code != 0:
        raise AfeError(code, load().afe_last_error().decode())
    return code


from .api import *  # noqa: E402,F401,F403
