"""pbfx — Python bindings of the gfx950 PBF solver core (libpbfx.so).

The compute path is the HIP library; this package only loads it, mirrors the
parameter structs and wraps the C ABI.  ``launch`` needs torch (device memory
and streams); ``params`` and ``sim`` do not.
"""
from . import params
from ._lib import LIB_PATH, PbfError, check, lib

__all__ = ["params", "lib", "check", "PbfError", "LIB_PATH"]
