"""MI355X-native L-BFGS-B engine: Python host layer over liblbfgsb_amd.so.

The C ABI (include/lbfgsb_amd.h) is the boundary; this package loads the
in-tree library with ctypes and uses PyTorch only for device memory and the
stream.  There is no CPU fallback: without the built library or a GPU the
calls raise.
"""
from . import problems  # noqa: F401
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["problems", "lib", "LIB_PATH"]
