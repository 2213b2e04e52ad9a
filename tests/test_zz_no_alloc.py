"""g6x_self_sweep_noalloc (include/g6x.h) through the C ABI: bitwise the
results of g6x_self_sweep on the same j-memory, on two calls in a row with
the second smaller (the static scratch keeps the first call's contents),
with outputs in caller-owned device memory and no change in the device's
free memory; and its refusals (j-memory not covering nj, not resident on
the device) as error returns."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.gpu_common import load_j, plummer_state

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime libg6.so links (already loaded with it), for the
    test's own device buffers.  torch ships a second copy of the runtime,
    which would not share libg6's device context."""
    global _hip
    if _hip is None:
        from starlab_amd import load
        load()
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip = L
    return _hip


def _dev(host):
    """A device copy of a host array (the test's allocation, not the library's)."""
    p = C.c_void_p()
    assert hip().hipMalloc(C.byref(p), host.nbytes) == 0
    assert hip().hipMemcpy(p, host.ctypes.data_as(C.c_void_p), host.nbytes, 1) == 0
    return p


def _host(p, like):
    out = np.empty_like(like)
    assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
    return out


def _free_bytes():
    f, t = C.c_size_t(), C.c_size_t()
    assert hip().hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def _noalloc(g, nj, i_first, ni, ti, eps2):
    fill = [np.full((ni, 3), np.nan), np.full((ni, 3), np.nan), np.full(ni, np.nan),
            np.full(ni, -7, np.int32)]
    bufs = [_dev(a) for a in fill]
    try:
        free0 = _free_bytes()
        rc = g.L.g6x_self_sweep_noalloc(g._cid.value, nj, i_first, ni, ti, eps2, *bufs)
        assert rc == 0, g.last_error()
        assert g.L.g6x_sync(g._cid.value) == 0
        free1 = _free_bytes()
        out = [_host(p, a) for p, a in zip(bufs, fill)]
    finally:
        for p in bufs:
            hip().hipFree(p)
    return (*out, free0, free1)


def test_noalloc_sweep_equals_self_sweep():
    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    assert g.open() == 0, g.last_error()
    try:
        cases = ((8192, 0, 8192, 0.0, 1e-8, 11), (3000, 100, 1000, 2.0 ** -12, 0.0, 12))
        for nj, i_first, ni, ti, eps2, seed in cases:
            print(f"case nj={nj} i_first={i_first} ni={ni} ti={ti} eps2={eps2}", flush=True)
            s = plummer_state(nj, seed=seed)
            load_j(g, s)
            # the allocating sweep places j-memory on the device and is the reference
            assert g.self_sweep(nj, i_first, ni, ti, eps2) == 0, g.last_error()
            ra, rj, rp, rn = g.read_results(ni)
            a, j, p, n, free0, free1 = _noalloc(g, nj, i_first, ni, ti, eps2)
            np.testing.assert_array_equal(a, ra)
            np.testing.assert_array_equal(j, rj)
            np.testing.assert_array_equal(p, rp)
            np.testing.assert_array_equal(n, rn)
            assert free1 == free0, (free0, free1)
            # and against the FP64 oracle on the predicted state
            xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], ti, 4)
            sel = np.arange(i_first, i_first + ni, 37)
            racc, _, rpot, rnn, _ = oracle.force(xp[sel], vp[sel], s["idx"][sel], xp, vp, s["m"],
                                                 s["idx"], eps2)
            assert np.array_equal(n[sel - i_first], rnn)
            assert np.max(np.abs(p[sel - i_first] - rpot) / np.abs(rpot)) <= 1e-6
        # j updates after the last placement go up without allocation: one
        # record (the staging buffer) and then all of them (one copy)
        s["x"][5] += 1e-3
        assert g.set_j_particle(5, 5, s["tj"][5], s["dtj"][5], s["m"][5], s["k18"][5],
                                s["j6"][5], s["a2"][5], s["v"][5], s["x"][5]) == 0
        a, j, p, n, free0, free1 = _noalloc(g, 3000, 0, 64, 0.0, 0.0)
        assert g.self_sweep(3000, 0, 64, 0.0, 0.0) == 0
        ra, rj, rp, rn = g.read_results(64)
        np.testing.assert_array_equal(a, ra)
        assert free1 == free0
        s["x"] += 1e-4
        load_j(g, s)
        a, j, p, n, free0, free1 = _noalloc(g, 3000, 0, 64, 0.0, 0.0)
        assert g.self_sweep(3000, 0, 64, 0.0, 0.0) == 0
        ra, rj, rp, rn = g.read_results(64)
        np.testing.assert_array_equal(a, ra)
        np.testing.assert_array_equal(n, rn)
        assert free1 == free0
    finally:
        g.close()


def test_noalloc_refusals():
    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    assert g.open() == 0, g.last_error()
    try:
        s = plummer_state(512, seed=2)
        load_j(g, s)
        # j-memory set on the host only: not resident, refused
        with pytest.raises(AssertionError, match="not resident"):
            _noalloc(g, 512, 0, 64, 0.0, 0.0)
        assert g.self_sweep(512, 0, 64, 0.0, 0.0) == 0
        # more j than were ever set: refused
        with pytest.raises(AssertionError, match="addresses"):
            _noalloc(g, 1 << 24, 0, 64, 0.0, 0.0)
        # and it still works afterwards
        _noalloc(g, 512, 0, 512, 0.0, 0.0)
    finally:
        g.close()
