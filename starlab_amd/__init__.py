"""starlab_amd -- MI355X-native GRAPE-6 force engine for kira (libg6.so).

The product is the C-ABI shared library ``starlab_amd/libg6.so`` (HIP for
gfx950), which exports the GRAPE-6 entry points of ``include/grape6.h`` and the
extensions of ``include/g6x.h``.  This package loads it with ctypes and offers
a Python mirror of the same interface (``starlab_amd.g6``).  There is no CPU
fallback: without the built library every entry raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBG6 = os.path.join(HERE, "libg6.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libg6.so (built by ``python -m starlab_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBG6):
            raise LibraryMissing(
                f"{LIBG6} is not built; run `python -m starlab_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        _lib = C.CDLL(LIBG6)
        from . import _abi
        _abi.declare(_lib)
    return _lib


def lib_path() -> str:
    return LIBG6
