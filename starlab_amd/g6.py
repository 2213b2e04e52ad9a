"""Python mirror of the GRAPE-6 ABI served by libg6.so.

Method names follow the C entry points of include/grape6.h (reference:
Starlab include/grape6.h:18-102) with the ``g6_``/``g6calc_`` prefix and the
trailing underscore dropped; arguments keep their meaning and order, with
numpy arrays in place of pointers.  Return codes are passed through unchanged
(0 = success), exactly as kira's driver sees them
(src/node/dyn/hdyn/evolve/hdyn_grape6.C).

The ``g6x``-prefixed helpers expose include/g6x.h (bulk j loads, the
device-resident sweep, timing); they have no counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import load

KERNELS = ("predict", "force", "reduce", "refine", "nblist", "scatter", "pot", "knn")


def _dp(a):
    return a.ctypes.data_as(C.c_void_p)


def _d3(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        a = a.reshape(n, 3)
    return a


class GRAPE6:
    """One process-wide libg6 device (cluster ids are accepted and ignored)."""

    def __init__(self, cluster_id: int = 0):
        self.L = load()
        self._cid = C.c_int(cluster_id)
        self._ref = C.byref(self._cid)

    # -- lifecycle ----------------------------------------------------------
    def open(self) -> int:
        return self.L.g6_open_(self._ref)

    def open_special(self, gpus) -> int:
        arr = (C.c_int * max(1, len(gpus)))(*gpus)
        return self.L.g6_open_special(len(gpus), arr)

    def close(self) -> int:
        return self.L.g6_close_(self._ref)

    def reset(self) -> int:
        return self.L.g6_reset_(self._ref)

    def reset_fofpga(self) -> int:
        return self.L.g6_reset_fofpga_(self._ref)

    def npipes(self) -> int:
        return self.L.g6_npipes_()

    # -- j-memory -------------------------------------------------------------
    def set_ti(self, ti: float) -> None:
        t = C.c_double(ti)
        self.L.g6_set_ti_(self._ref, C.byref(t))

    def set_j_particle(self, address, index, tj, dtj, mass, k18, j6, a2, v, x) -> int:
        vecs = [np.ascontiguousarray(q, dtype=np.float64) for q in (k18, j6, a2, v, x)]
        return self.L.g6_set_j_particle_(
            self._ref, C.byref(C.c_int(address)), C.byref(C.c_int(index)),
            C.byref(C.c_double(tj)), C.byref(C.c_double(dtj)), C.byref(C.c_double(mass)),
            *[_dp(q) for q in vecs])

    def set_j_particles(self, address, index, tj, dtj, mass, k18, j6, a2, v, x) -> int:
        """Bulk load (g6x_set_j_particles)."""
        n = len(mass)
        addr = np.ascontiguousarray(address, dtype=np.int32)
        idx = np.ascontiguousarray(index, dtype=np.int32)
        sc = [np.ascontiguousarray(q, dtype=np.float64) for q in (tj, dtj, mass)]
        vecs = [_d3(q, n) for q in (k18, j6, a2, v, x)]
        return self.L.g6x_set_j_particles(self._cid.value, n, _dp(addr), _dp(idx),
                                          *[_dp(q) for q in sc], *[_dp(q) for q in vecs])

    # -- force ------------------------------------------------------------------
    def calc_firsthalf(self, nj, index, xi, vi, eps2, h2, aold=None, j6old=None,
                       phiold=None) -> None:
        ni = len(index)
        self._idx = np.ascontiguousarray(index, dtype=np.int32)
        self._xi = _d3(xi, ni)
        self._vi = _d3(vi, ni)
        self._h2 = np.ascontiguousarray(h2, dtype=np.float64)
        self._eps2 = C.c_double(eps2)
        z3 = np.zeros((ni, 3))
        a = _d3(aold, ni) if aold is not None else z3
        j = _d3(j6old, ni) if j6old is not None else z3
        p = np.ascontiguousarray(phiold if phiold is not None else np.zeros(ni), dtype=np.float64)
        self._keep = (a, j, p)
        self.L.g6calc_firsthalf_(self._ref, C.byref(C.c_int(nj)), C.byref(C.c_int(ni)),
                                 _dp(self._idx), _dp(self._xi), _dp(self._vi), _dp(a), _dp(j),
                                 _dp(p), C.byref(self._eps2), _dp(self._h2))

    def calc_lasthalf2(self, nj, ni):
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        nn = np.zeros(ni, np.int32)
        rc = self.L.g6calc_lasthalf2_(self._ref, C.byref(C.c_int(nj)), C.byref(C.c_int(ni)),
                                      _dp(self._idx), _dp(self._xi), _dp(self._vi),
                                      C.byref(self._eps2), _dp(self._h2), _dp(acc), _dp(jerk),
                                      _dp(pot), _dp(nn))
        return rc, acc, jerk, pot, nn

    def calc_lasthalf(self, nj, ni):
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        rc = self.L.g6calc_lasthalf_(self._ref, C.byref(C.c_int(nj)), C.byref(C.c_int(ni)),
                                     _dp(self._idx), _dp(self._xi), _dp(self._vi),
                                     C.byref(self._eps2), _dp(self._h2), _dp(acc), _dp(jerk),
                                     _dp(pot))
        return rc, acc, jerk, pot

    def read_neighbour_list(self) -> int:
        return self.L.g6_read_neighbour_list_(self._ref)

    def get_neighbour_list(self, pipe: int, max_length: int):
        buf = np.zeros(max(max_length, 1), np.int32)
        nblen = C.c_int(0)
        rc = self.L.g6_get_neighbour_list_(self._ref, C.byref(C.c_int(pipe)),
                                           C.byref(C.c_int(max_length)), C.byref(nblen),
                                           _dp(buf))
        return rc, buf[:max(0, min(nblen.value, max_length))].copy()

    def force(self, nj, index, xi, vi, eps2, h2=None):
        """firsthalf + lasthalf2 over i-blocks of at most npipes (as kira chunks)."""
        ni = len(index)
        h2 = np.zeros(ni) if h2 is None else np.asarray(h2, dtype=np.float64)
        xi = _d3(xi, ni); vi = _d3(vi, ni)
        np_ = self.npipes()
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        nn = np.zeros(ni, np.int32)
        for s in range(0, ni, np_):
            e = min(ni, s + np_)
            self.calc_firsthalf(nj, index[s:e], xi[s:e], vi[s:e], eps2, h2[s:e])
            rc, a, j, p, n = self.calc_lasthalf2(nj, e - s)
            if rc:
                raise RuntimeError(f"g6calc_lasthalf2_ returned {rc}: {self.last_error()}")
            acc[s:e], jerk[s:e], pot[s:e], nn[s:e] = a, j, p, n
        return acc, jerk, pot, nn

    # -- extensions -------------------------------------------------------------
    def self_sweep(self, nj, i_first, ni, ti, eps2) -> int:
        return self.L.g6x_self_sweep(self._cid.value, nj, i_first, ni, ti, eps2)

    def potential(self, nj, i_first, ni, ti, eps2):
        """pot of j-particles [i_first, i_first+ni) against [0, nj) at ti
        (include/g6x.h g6x_potential)."""
        pot = np.zeros(ni)
        rc = self.L.g6x_potential(self._cid.value, nj, i_first, ni, ti, eps2, _dp(pot))
        if rc:
            raise RuntimeError(f"g6x_potential: {self.last_error()}")
        return pot

    def densities(self, nj, i_first, ni, ti, k=12):
        """(density, dk2) of j-particles [i_first, i_first+ni) among [0, nj)
        (include/g6x.h g6x_densities; compute_density with k neighbours)."""
        rho = np.zeros(ni); dk2 = np.zeros(ni)
        rc = self.L.g6x_densities(self._cid.value, nj, i_first, ni, ti, k, _dp(rho), _dp(dk2))
        if rc:
            raise RuntimeError(f"g6x_densities: {self.last_error()}")
        return rho, dk2

    def read_results(self, ni):
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        nn = np.zeros(ni, np.int32)
        rc = self.L.g6x_read_results(self._cid.value, ni, _dp(acc), _dp(jerk), _dp(pot), _dp(nn))
        if rc:
            raise RuntimeError(f"g6x_read_results: {self.last_error()}")
        return acc, jerk, pot, nn

    def read_predicted(self, nj):
        xp = np.zeros((nj, 3)); vp = np.zeros((nj, 3))
        rc = self.L.g6x_read_predicted(self._cid.value, nj, _dp(xp), _dp(vp))
        if rc:
            raise RuntimeError(f"g6x_read_predicted: {self.last_error()}")
        return xp, vp

    def sync(self) -> int:
        return self.L.g6x_sync(self._cid.value)

    def option(self, key: str, value=None):
        if value is None:
            return self.L.g6x_get_option(key.encode())
        rc = self.L.g6x_set_option(key.encode(), int(value))
        if rc:
            raise KeyError(key)
        return value

    def kernel_time(self, name: str):
        ms = C.c_double(0)
        n = self.L.g6x_kernel_time(KERNELS.index(name), C.byref(ms))
        return n, ms.value

    def reset_timing(self) -> None:
        self.L.g6x_reset_timing()

    def geometry(self):
        wg, ns, jc = C.c_int(), C.c_int(), C.c_int()
        self.L.g6x_last_force_geometry(C.byref(wg), C.byref(ns), C.byref(jc))
        return wg.value, ns.value, jc.value

    def refinements(self) -> int:
        return self.L.g6x_last_refinements()

    def autotune(self, nj: int, ni_max: int, eps2: float = 0.0) -> int:
        n = self.L.g6x_autotune(0, int(nj), int(ni_max), float(eps2))
        if n < 0:
            raise RuntimeError(self.last_error())
        return n

    def tuning(self):
        """Tuned shapes: {bucket: (variant, tile, waves, lanes)}."""
        out = {}
        for b in range(21):
            v, t, w, l = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            if self.L.g6x_get_tuning(b, C.byref(v), C.byref(t), C.byref(w), C.byref(l)) == 1:
                out[b] = (v.value, t.value, w.value, l.value)
        return out

    def clear_tuning(self):
        self.L.g6x_clear_tuning()

    def save_tuning(self, path: str) -> int:
        return self.L.g6x_save_tuning(str(path).encode())

    def load_tuning(self, path: str) -> int:
        return self.L.g6x_load_tuning(str(path).encode())

    def last_error(self) -> str:
        s = self.L.g6x_last_error()
        return s.decode() if s else ""
