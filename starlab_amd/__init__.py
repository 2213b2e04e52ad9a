"""starlab_amd -- MI355X-native GRAPE force engine for kira (libg6.so, libg4x.so).

The product is the C-ABI shared library ``starlab_amd/libg6.so`` (HIP for
gfx950), which exports the GRAPE-6 entry points of ``include/grape6.h`` and the
extensions of ``include/g6x.h``, and its sibling ``starlab_amd/libg4x.so``,
which exports the GRAPE-4 entry points of ``include/grape4.h`` on the same
kernels (``starlab_amd.g4``).  This package loads it with ctypes and offers
a Python mirror of the same interface (``starlab_amd.g6``).  There is no CPU
fallback: without the built library every entry raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBG6 = os.path.join(HERE, "libg6.so")
LIBG4X = os.path.join(HERE, "libg4x.so")

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


_lib_g4 = None


def load_g4x() -> C.CDLL:
    """Load libg4x.so, the GRAPE-4 ABI on the same kernels (built with libg6)."""
    global _lib_g4
    if _lib_g4 is None:
        if not os.path.exists(LIBG4X):
            raise LibraryMissing(
                f"{LIBG4X} is not built; run `python -m starlab_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        _lib_g4 = C.CDLL(LIBG4X)
        from . import _abi
        _abi.declare_g4(_lib_g4)
    return _lib_g4
