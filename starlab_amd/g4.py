"""Python mirror of the GRAPE-4 ABI served by libg4x.so.

Method names follow the C entry points of include/grape4.h (reference:
Starlab include/grape4.h:22-45) with the ``h3`` prefix and the trailing
underscore dropped; arguments keep their meaning and order, with numpy
arrays in place of pointers.  The calls are those kira's GRAPE-4 driver
makes (src/node/dyn/hdyn/evolve/hdyn_grape4.C): ``mjpdma_indirect`` +
``mjpdma_start`` to load j, ``setti``, ``calc_firsthalf``, ``wait``,
``calc_lasthalf``, ``nbread`` and ``nblist``.

The GRAPE-4 calls return no status.  By default the library exits the
process on a failure; ``GRAPE4(exit_on_error=False)`` keeps the process
alive and ``errors()`` / ``last_error()`` report what failed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import load_g4x


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _d3(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


class GRAPE4:
    """The process-wide libg4x device."""

    def __init__(self, exit_on_error: bool = True):
        self.L = load_g4x()
        self.L.g4x_set_option(b"exit_on_error", 1 if exit_on_error else 0)

    # -- lifecycle / configuration ------------------------------------------
    def open(self) -> None:
        self.L.h3open_()

    def close(self) -> None:
        self.L.h3close_()

    def npipe(self) -> int:
        return self.L.h3npipe_()

    def jpmax(self) -> int:
        return self.L.h3jpmax_()

    def getnboards(self) -> int:
        return self.L.h3getnboards_()

    def wait(self) -> int:
        return self.L.h3wait_()

    def setti(self, ti: float) -> None:
        self.L.h3setti_(C.byref(C.c_double(ti)))

    # -- j-memory ---------------------------------------------------------------
    def _dma_args(self, hostindex, xj, vj, aj, jj, mj, tj):
        arrs = (np.ascontiguousarray(hostindex, dtype=np.int32), _d3(xj), _d3(vj), _d3(aj),
                _d3(jj), _d(mj), _d(tj))
        self._keep = arrs
        return arrs

    def jpdma_indirect(self, hostindex, xj, vj, aj, jj, mj, tj) -> None:
        """hostindex: 1-based indices into the caller arrays (address = index - 1)."""
        h, x, v, a, j, m, t = self._dma_args(hostindex, xj, vj, aj, jj, mj, tj)
        self.L.h3jpdma_indirect_(C.byref(C.c_int(len(h))), _p(h), _p(x), _p(v), _p(a), _p(j),
                                 _p(m), _p(t), C.byref(C.c_int(1)))

    def mjpdma_indirect(self, hostindex, xj, vj, aj, jj, mj, tj, buff_id: int) -> None:
        h, x, v, a, j, m, t = self._dma_args(hostindex, xj, vj, aj, jj, mj, tj)
        self.L.h3mjpdma_indirect_(C.byref(C.c_int(len(h))), _p(h), _p(x), _p(v), _p(a),
                                  _p(j), _p(m), _p(t), C.byref(C.c_int(1)),
                                  C.byref(C.c_int(buff_id)))

    def mjpdma_start(self, buff_id: int) -> None:
        self.L.h3mjpdma_start_(C.byref(C.c_int(buff_id)))

    def mjpdma_flush(self) -> None:
        self.L.h3mjpdma_flush_()

    def load_all(self, x, v, a2, j6, m, tj, batch: int | None = None) -> None:
        """Load particles 0..n-1 the way kira's jpdma_nodes does
        (hdyn_grape4.C:240-295): batches of h3jpmax_() through buffers 0..5."""
        n = len(m)
        batch = batch or self.jpmax()
        buf = 0
        for s in range(0, n, batch):
            idx = np.arange(s + 1, min(n, s + batch) + 1, dtype=np.int32)
            self.mjpdma_indirect(idx, x, v, a2, j6, m, tj, buf)
            self.mjpdma_start(buf)
            buf = 0 if buf >= 5 else buf + 1

    # -- force ------------------------------------------------------------------
    def calc_firsthalf(self, nj: int, xi, vi, eps2, h2) -> None:
        xi = _d3(xi)
        ni = len(xi)
        self._i = (xi, _d3(vi), _d(np.broadcast_to(np.asarray(eps2, float), (ni,))),
                   _d(np.broadcast_to(np.asarray(h2, float), (ni,))))
        x, v, e, h = self._i
        self.L.h3calc_firsthalf_(C.byref(C.c_int(nj)), C.byref(C.c_int(ni)), _p(x), _p(v),
                                 _p(e), _p(h))

    def calc_lasthalf(self, ni: int):
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        self.L.h3calc_lasthalf_(C.byref(C.c_int(ni)), _p(acc), _p(jerk), _p(pot))
        return acc, jerk, pot

    def calc(self, nj: int, xi, vi, eps2, h2):
        xi = _d3(xi)
        ni = len(xi)
        x, v = xi, _d3(vi)
        e = _d(np.broadcast_to(np.asarray(eps2, float), (ni,)))
        h = _d(np.broadcast_to(np.asarray(h2, float), (ni,)))
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        self.L.h3calc_(C.byref(C.c_int(nj)), C.byref(C.c_int(ni)), _p(x), _p(v), _p(e), _p(h),
                       _p(acc), _p(jerk), _p(pot))
        return acc, jerk, pot

    # -- neighbour lists --------------------------------------------------------
    def nbread(self) -> None:
        self.L.h3nbread_(C.byref(C.c_int(self.getnboards())))

    def nblist(self, chip: int, cap: int) -> np.ndarray:
        """List of pipe `chip`; `cap` must be at least its length (<= nj)."""
        buf = np.zeros(max(cap, 1), np.int32)
        n = self.L.h3nblist_(C.byref(C.c_int(0)), C.byref(C.c_int(chip)), _p(buf))
        return buf[:n].copy()

    def count_nblist(self, chip: int) -> int:
        return self.L.count_nblist_low(0, chip)

    # -- extensions -------------------------------------------------------------
    def errors(self) -> int:
        return self.L.g4x_errors()

    def last_error(self) -> str:
        s = self.L.g4x_last_error()
        return s.decode() if s else ""

    def option(self, key: str, value=None) -> int:
        if value is None:
            v = self.L.g4x_get_option(key.encode())
            if v == -1 and key not in ("exit_on_error",):
                raise KeyError(key)
            return v
        if self.L.g4x_set_option(key.encode(), int(value)) != 0:
            raise KeyError(key)
        return int(value)
