"""CPU oracle for the GRAPE-6 force path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only to check results.  The product
(``starlab_amd`` / ``libg6.so``) never imports, links or calls it.

``liboracle_g6.so`` is built from ``g6_oracle.c`` (FP64, no FMA contraction)
by ``oracle/Makefile``.  ``oracle/_ref/libharp3sim.so`` is the reference's
own IEEE GRAPE-4 simulator (``src/harpsims/harp3fastsim.c``) compiled from
its source; it exists only where the reference tree was present at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle_g6.so")
_REF = os.path.join(HERE, "_ref", "libharp3sim.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def host_cpu():
    """(model name, has AVX-512F) of this host, from /proc/cpuinfo."""
    model, avx512 = "unknown", False
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags"):
                    avx512 = " avx512f " in f" {line.split(':', 1)[1].strip()} "
                    break
    except OSError:
        pass
    return model, avx512


_fast = None


def fast_lib():
    """The speed build of the same source for the benchmark's CPU baseline
    (-O3, FMA contraction, OpenMP; x86-64-v4 on AVX-512 hosts, else
    x86-64-v3): (CDLL, compiler flags).  Not a checker."""
    global _fast
    if _fast is None:
        _, avx512 = host_cpu()
        level = "v4" if avx512 else "v3"
        path = os.path.join(HERE, f"liboracle_g6_{level}.so")
        if not os.path.exists(path):
            build()
        _fast = (_prototypes(C.CDLL(path)), f"gcc -O3 -march=x86-64-{level} -fopenmp")
    return _fast


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _prototypes(C.CDLL(_LIB))
    return _lib


def _prototypes(L):
    L.or_ran2_seed.argtypes = [C.c_int]
    L.or_ran2_seed.restype = C.c_int
    L.or_randinter.argtypes = [C.c_double, C.c_double]
    L.or_randinter.restype = C.c_double
    L.or_n_rand.restype = C.c_int
    L.or_threads.restype = C.c_int
    L.or_makeplummer.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_predict.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                             C.c_double, C.c_int, _dp, _dp]
    L.or_force.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp,
                           _ip, C.c_double, _dp, _dp, _dp, _ip, _dp]
    L.or_force_ld.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp,
                              _ip, C.c_double, _dp, _dp, _dp]
    L.or_neighbours.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, C.c_int,
                                _dp, _ip, C.c_int, _ip, _ip]
    L.or_g4_force.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp,
                              _dp, _dp, _dp, C.c_int, _ip, _ip]
    L.or_density.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.or_energies.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double,
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_hermite_shared.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_longlong,
                                    C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_hermite_shared.restype = C.c_int
    L.or_hermite_block.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_longlong,
                                   C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_longlong), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    L.or_hermite_block.restype = C.c_int
    return L


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# -- random numbers / initial conditions ------------------------------------

def ran2_sequence(seed: int, n: int, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    L = lib()
    L.or_ran2_seed(seed)
    return np.array([L.or_randinter(a, b) for _ in range(n)])


def makeplummer(n: int, seed: int, scale: bool = True):
    m = np.zeros(n)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    lib().or_makeplummer(n, seed, int(scale), m, x, v)
    return m, x, v


# -- predictor / force / neighbours / energies ------------------------------

def predict(x, v, a2, j6, k18, tj, ti: float, order: int = 4):
    x = _d(x); nj = x.shape[0]
    xp = np.zeros((nj, 3)); vp = np.zeros((nj, 3))
    lib().or_predict(nj, x, _d(v), _d(a2), _d(j6), _d(k18), _d(tj), ti, order, xp, vp)
    return xp, vp


def force(xi, vi, excl, xj, vj, mj, jindex, eps2: float):
    xi = _d(xi); ni = xi.shape[0]; nj = _d(xj).shape[0]
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    nn = np.zeros(ni, np.int32); dnn2 = np.zeros(ni)
    lib().or_force(ni, xi, _d(vi), _i(excl), nj, _d(xj), _d(vj), _d(mj),
                   _i(jindex), eps2, acc, jerk, pot, nn, dnn2)
    return acc, jerk, pot, nn, dnn2


def force_ld(xi, vi, excl: int, xj, vj, mj, jindex, eps2: float):
    acc = np.zeros(3); jerk = np.zeros(3); pot = np.zeros(1)
    lib().or_force_ld(_d(xi), _d(vi), int(excl), _d(xj).shape[0], _d(xj), _d(vj),
                      _d(mj), _i(jindex), eps2, acc, jerk, pot)
    return acc, jerk, pot[0]


def neighbours(xi, h2, excl, include_self: bool, xj, jindex, cap: int):
    xi = _d(xi); ni = xi.shape[0]
    counts = np.zeros(ni, np.int32)
    lists = np.full(ni * max(cap, 1), -1, np.int32)
    lib().or_neighbours(ni, xi, _d(h2), _i(excl), int(include_self), _d(xj).shape[0],
                        _d(xj), _i(jindex), cap, counts, lists)
    return counts, lists.reshape(ni, max(cap, 1))


def g4_force(xi, vi, eps2, h2, xj, vj, mj, cap: int = 1024):
    """GRAPE-4 semantics (reference simulator pair loop, pot added): returns
    acc, jerk, pot, counts, lists[ni, cap] (j addresses, -1 padded)."""
    xi = _d(xi).reshape(-1, 3); ni = xi.shape[0]; nj = _d(xj).reshape(-1, 3).shape[0]
    eps2 = _d(np.broadcast_to(np.asarray(eps2, dtype=np.float64), (ni,)))
    h2 = _d(np.broadcast_to(np.asarray(h2, dtype=np.float64), (ni,)))
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    counts = np.zeros(ni, np.int32)
    lists = np.full(ni * max(cap, 1), -1, np.int32)
    lib().or_g4_force(ni, xi, _d(vi), eps2, h2, nj, _d(xj), _d(vj), _d(mj), acc, jerk, pot,
                      max(cap, 1), counts, lists)
    return acc, jerk, pot, counts, lists.reshape(ni, max(cap, 1))


def density(m, x, k: int = 12, i_first: int = 0, ni: int | None = None):
    """compute_density (whole-system mode): (density, dk2) of particles
    i_first .. i_first+ni-1."""
    x = _d(x).reshape(-1, 3); n = x.shape[0]
    ni = n - i_first if ni is None else ni
    rho = np.zeros(ni); dk2 = np.zeros(ni)
    lib().or_density(n, _d(m), x, i_first, ni, k, rho, dk2)
    return rho, dk2


def hermite_shared(m, x, v, t_end, eta, eps2, dt_max=0.125, max_steps=10**8, L=None):
    """Integrate one system with the shared-step Hermite rules; returns
    (x, v, t_final, steps, e0, e1, rc).  L: another build of the same source
    (fast_lib()[0], for the CPU baseline); default the oracle."""
    x = np.array(x, dtype=np.float64, order="C")
    v = np.array(v, dtype=np.float64, order="C")
    t = C.c_double(); ns = C.c_longlong(); e0 = C.c_double(); e1 = C.c_double()
    rc = (L or lib()).or_hermite_shared(len(m), _d(m), x, v, t_end, eta, eps2, dt_max,
                                        max_steps, C.byref(t), C.byref(ns), C.byref(e0),
                                        C.byref(e1))
    return x, v, t.value, ns.value, e0.value, e1.value, rc


def hermite_block(m, x, v, t_end, eta, eps2, dt_max=0.125, max_blocks=10**8, L=None):
    """Integrate one system with kira's block-step Hermite rules; returns
    (x, v, t_final, blocks, particle_steps, e0, e1, rc)."""
    x = np.array(x, dtype=np.float64, order="C")
    v = np.array(v, dtype=np.float64, order="C")
    t = C.c_double(); nb = C.c_longlong(); ps = C.c_longlong()
    e0 = C.c_double(); e1 = C.c_double()
    rc = (L or lib()).or_hermite_block(len(m), _d(m), x, v, t_end, eta, eps2, dt_max,
                                       max_blocks, C.byref(t), C.byref(nb), C.byref(ps),
                                       C.byref(e0), C.byref(e1))
    return x, v, t.value, nb.value, ps.value, e0.value, e1.value, rc


def threads() -> int:
    return lib().or_threads()


def energies(m, x, v, eps2: float = 0.0):
    ep = C.c_double(); ek = C.c_double()
    lib().or_energies(len(m), _d(m), _d(x), _d(v), eps2, C.byref(ep), C.byref(ek))
    return ep.value, ek.value


# -- the reference's own GRAPE-4 IEEE simulator (oracle/_ref) ----------------

class Harp3Sim:
    """ctypes view of ``harp3fastsim.c`` (GRAPE-4 ABI; 3 pipes, <= 1000 j).

    Conventions of that simulator (harp3fastsim.c:223-322): eps2 and h2 are
    doubled internally, the neighbour test is on r^2 + 2 eps2 < 2 h2, self is
    skipped by r^2 > 0, and pot is not computed.
    """

    NMAX, IPMAX, JPMAX, NNBMAX = 1000, 3, 26, 1024

    def __init__(self):
        if not os.path.exists(_REF):
            raise FileNotFoundError(_REF)
        L = C.CDLL(_REF)
        self.L = L
        L.h3setti_.argtypes = [C.POINTER(C.c_double)]
        L.h3jpdma_indirect_.argtypes = [C.POINTER(C.c_int), _ip, _dp, _dp, _dp,
                                        _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.h3calc_firsthalf_.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        _dp, _dp, _dp, _dp]
        L.h3calc_lasthalf_.argtypes = [C.POINTER(C.c_int), _dp, _dp, _dp]
        L.h3nblist_.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), _ip]
        L.h3nblist_.restype = C.c_int

    @staticmethod
    def available() -> bool:
        return os.path.exists(_REF)

    def set_j(self, x, v, a2, j6, m, tj):
        """Load all j (hostindex is 1-based, at most JPMAX per DMA call)."""
        n = len(m)
        assert n <= self.NMAX
        x, v, a2, j6, m, tj = (_d(t) for t in (x, v, a2, j6, m, tj))
        mode = C.c_int(0)
        for s in range(0, n, self.JPMAX):
            cnt = min(self.JPMAX, n - s)
            idx = _i(np.arange(s + 1, s + cnt + 1))
            self.L.h3jpdma_indirect_(C.byref(C.c_int(cnt)), idx, x, v, a2, j6, m, tj,
                                     C.byref(mode))

    def predicted(self, nj):
        xp = np.ctypeslib.as_array((C.c_double * (self.NMAX * 3)).in_dll(self.L, "xjpred"))
        vp = np.ctypeslib.as_array((C.c_double * (self.NMAX * 3)).in_dll(self.L, "vjpred"))
        return xp.reshape(self.NMAX, 3)[:nj].copy(), vp.reshape(self.NMAX, 3)[:nj].copy()

    def calc(self, ti, nj, xi, vi, eps2, h2):
        """Force on <= 3 i-particles; returns acc, jerk, neighbour lists."""
        ni = len(xi)
        assert ni <= self.IPMAX
        t = C.c_double(ti)
        self.L.h3setti_(C.byref(t))
        e = _d(np.full(ni, eps2 / 2.0))
        h = _d(np.asarray(h2, dtype=np.float64) / 2.0)
        self.L.h3calc_firsthalf_(C.byref(C.c_int(nj)), C.byref(C.c_int(ni)),
                                 _d(xi), _d(vi), e, h)
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        self.L.h3calc_lasthalf_(C.byref(C.c_int(ni)), acc, jerk, pot)
        lists = []
        for i in range(ni):
            buf = np.zeros(self.NNBMAX, np.int32)
            cnt = self.L.h3nblist_(C.byref(C.c_int(0)), C.byref(C.c_int(i)), buf)
            lists.append(buf[:cnt].copy())
        return acc, jerk, lists
