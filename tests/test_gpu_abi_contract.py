"""The GRAPE-6 ABI's contract on misuse and lifecycle (include/grape6.h),
through the C entry points: errors come back as return codes (kira's
retry/abort logic reads them, hdyn_grape6.C:744-754, 2087-2166) and never as a
device fault or a stale result."""
import numpy as np
import pytest

import oracle
from tests.gpu_common import compare, load_j, plummer_state

pytestmark = pytest.mark.gpu


@pytest.fixture()
def g():
    from starlab_amd.g6 import GRAPE6
    dev = GRAPE6()
    assert dev.open() == 0, dev.last_error()
    yield dev
    dev.close()


def _block(g, s, isel, eps2=1e-8):
    nj = len(s["m"])
    g.set_ti(0.0)
    xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], 0.0, 4)
    idx = s["idx"][isel]
    acc, jerk, pot, nn = g.force(nj, idx, xp[isel], vp[isel], eps2)
    ref = oracle.force(xp[isel], vp[isel], idx, xp, vp, s["m"], s["idx"], eps2)
    return compare(acc, jerk, pot, nn, ref, xp[isel], vp[isel], s["m"], xp, vp, eps2,
                   np.asarray(isel))


def test_oversized_block_is_an_error_then_recovers(g):
    s = plummer_state(3000, seed=51)
    load_j(g, s)
    npipes = g.npipes()
    ni = npipes + 1
    idx = np.arange(ni, dtype=np.int32) % 3000
    z = np.zeros((ni, 3))
    g.calc_firsthalf(3000, idx, z, z, 1e-8, np.zeros(ni))
    assert "npipes" in g.last_error()
    rc, *_ = g.calc_lasthalf2(3000, ni)
    assert rc != 0
    errs = _block(g, s, np.arange(100))
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6


def test_negative_address_is_rejected(g):
    z = np.zeros(3)
    assert g.set_j_particle(-1, 0, 0.0, 1.0, 1.0, z, z, z, z, z) != 0


def test_close_reopen_drops_the_last_calculation(g):
    """After g6_close_ / g6_open_ (kira's release and re-attach) nothing of
    the last block is left to read: lasthalf and the list calls return
    errors; a new block after reloading j works."""
    s = plummer_state(2000, seed=53)
    load_j(g, s)
    _block(g, s, np.arange(64))
    assert g.close() == 0
    assert g.open() == 0
    rc, *_ = g.calc_lasthalf2(2000, 64)
    assert rc != 0
    assert g.read_neighbour_list() < 0
    rc, lst = g.get_neighbour_list(0, 16)
    assert rc != 0 and len(lst) == 0
    load_j(g, s)
    errs = _block(g, s, np.arange(64))
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6


def test_neighbour_list_pipe_out_of_range(g):
    s = plummer_state(1000, seed=55)
    load_j(g, s)
    g.set_ti(0.0)
    xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], 0.0, 4)
    idx = np.arange(8, dtype=np.int32)
    g.calc_firsthalf(1000, idx, xp[:8], vp[:8], 1e-8, np.full(8, 0.01))
    rc, *_ = g.calc_lasthalf2(1000, 8)
    assert rc == 0
    assert g.read_neighbour_list() == 0
    for pipe in (-1, 8, 1 << 20):
        rc, lst = g.get_neighbour_list(pipe, 16)
        assert rc != 0 and len(lst) == 0


def test_unused_entry_points_are_harmless(g):
    """The unit setters and the j-DMA buffer calls exist for linking and do
    nothing (kira never calls them; GRAPE-6 fixed-point units have no meaning
    here)."""
    import ctypes as C
    L = g.L
    one = C.c_int(1)
    L.g6_set_tunit_(C.byref(one))
    L.g6_set_xunit_(C.byref(one))
    assert L.g6_initialize_jp_buffer_(C.byref(C.c_int(0)), C.byref(C.c_int(256))) == 0
    assert L.g6_flush_jp_buffer_(C.byref(C.c_int(0))) == 0
    s = plummer_state(500, seed=57)
    load_j(g, s)
    errs = _block(g, s, np.arange(50))
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6


def test_read_results_only_after_a_self_sweep(g):
    """g6x_read_results returns the last g6x_self_sweep's results, and an
    error once an ABI block or another sweep has used the device since."""
    s = plummer_state(2048, seed=59)
    load_j(g, s)
    assert g.self_sweep(2048, 0, 2048, 0.0, 1e-8) == 0
    acc, _, pot, _ = g.read_results(2048)
    assert np.all(pot < 0) and np.isfinite(acc).all()
    _block(g, s, np.arange(32))
    with pytest.raises(RuntimeError, match="self_sweep"):
        g.read_results(2048)
    g.potential(2048, 0, 2048, 0.0, 1e-8)
    with pytest.raises(RuntimeError, match="self_sweep"):
        g.read_results(2048)
