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


def _raw_block(g, nj, index, xi, vi, h2, eps2=1e-8):
    """g6calc_firsthalf_ / g6calc_lasthalf2_ through the raw C entry points,
    so that NULL arrays reach the library as a C caller would pass them."""
    import ctypes as C
    ni = next(len(a) for a in (index, xi, vi, h2) if a is not None)
    cid, cnj, cni = C.c_int(0), C.c_int(nj), C.c_int(ni)
    e = C.c_double(eps2)
    z3 = np.zeros((ni, 3))
    zp = np.zeros(ni)
    acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
    nn = np.zeros(ni, np.int32)
    ptr = lambda a: None if a is None else a.ctypes.data   # noqa: E731
    args = (ptr(index), ptr(xi), ptr(vi))
    g.L.g6calc_firsthalf_(C.byref(cid), C.byref(cnj), C.byref(cni), *args, z3.ctypes.data,
                          z3.ctypes.data, zp.ctypes.data, C.byref(e), ptr(h2))
    first_error = g.last_error()
    rc = g.L.g6calc_lasthalf2_(C.byref(cid), C.byref(cnj), C.byref(cni), *args, C.byref(e),
                               ptr(h2), acc.ctypes.data, jerk.ctypes.data, pot.ctypes.data,
                               nn.ctypes.data)
    return rc, acc, jerk, pot, nn, first_error


def test_null_arrays_from_a_c_caller(g):
    """A NULL h2 means no neighbour spheres (forces as usual, empty lists);
    NULL indices, positions or velocities are an error, not a crash."""
    n, ni = 1500, 32
    s = plummer_state(n, seed=61)
    load_j(g, s)
    g.set_ti(0.0)
    xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], 0.0, 4)
    isel = np.arange(ni)
    idx = np.ascontiguousarray(s["idx"][isel])
    xi = np.ascontiguousarray(xp[isel])
    vi = np.ascontiguousarray(vp[isel])
    rc, acc, jerk, pot, nn, _ = _raw_block(g, n, idx, xi, vi, None)
    assert rc == 0, g.last_error()
    ref = oracle.force(xi, vi, idx, xp, vp, s["m"], s["idx"], 1e-8)
    errs = compare(acc, jerk, pot, nn, ref, xi, vi, s["m"], xp, vp, 1e-8, isel)
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6
    assert g.read_neighbour_list() == 0
    for pipe in (0, ni - 1):
        rc, lst = g.get_neighbour_list(pipe, 16)
        assert rc == 0 and len(lst) == 0
    for missing in range(3):
        arrs = [idx, xi, vi]
        arrs[missing] = None
        rc, *_, first_error = _raw_block(g, n, *arrs, np.zeros(ni))
        assert rc != 0 and "null" in first_error
    errs = _block(g, s, np.arange(64))
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6


def test_autotune_drops_lists_and_sweep_results(g):
    """g6x_autotune runs its candidates on the i-block and results buffers:
    afterwards the last block's neighbour lists and the last self-sweep's
    results report errors instead of returning the tuner's data."""
    n = 2048
    s = plummer_state(n, seed=63)
    load_j(g, s)
    try:
        assert g.self_sweep(n, 0, n, 0.0, 1e-8) == 0
        g.autotune(n, 64, 1e-8)
        with pytest.raises(RuntimeError, match="self_sweep"):
            g.read_results(n)
        g.set_ti(0.0)
        xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], 0.0, 4)
        idx = np.arange(8, dtype=np.int32)
        g.calc_firsthalf(n, idx, xp[:8], vp[:8], 1e-8, np.full(8, 0.01))
        rc, *_ = g.calc_lasthalf2(n, 8)
        assert rc == 0
        g.autotune(n, 64, 1e-8)
        assert g.read_neighbour_list() < 0
        rc, lst = g.get_neighbour_list(0, 16)
        assert rc != 0 and len(lst) == 0
    finally:
        g.clear_tuning()
    errs = _block(g, s, np.arange(64))
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6


def test_extension_argument_errors(g):
    """Negative counts and NULL arrays of the extension entry points are
    errors; a failed self-sweep leaves no results to read."""
    import ctypes as C
    L = g.L
    s = plummer_state(600, seed=65)
    load_j(g, s)
    assert L.g6x_set_j_particles(0, -1, None, None, None, None, None, None, None, None,
                                 None, None) != 0
    assert L.g6x_set_j_particles(0, 1, None, None, None, None, None, None, None, None,
                                 None, None) != 0
    assert L.g6x_set_j_particles(0, 0, None, None, None, None, None, None, None, None,
                                 None, None) == 0
    assert g.self_sweep(600, 0, 600, 0.0, 1e-8) == 0
    buf = np.zeros(3 * 600)
    assert L.g6x_read_results(0, -1, buf.ctypes.data, None, None, None) != 0
    assert L.g6x_read_predicted(0, -5, buf.ctypes.data, None) != 0
    assert g.self_sweep(600, 0, 601, 0.0, 1e-8) != 0      # range outside j-memory
    with pytest.raises(RuntimeError, match="self_sweep"):
        g.read_results(600)


@pytest.mark.parametrize("pre_dirty", [100, 1500], ids=["scatter", "full_upload"])
def test_set_j_between_halves_keeps_the_block_in_flight(g, pre_dirty):
    """j records sent between g6calc_firsthalf_ and g6calc_lasthalf2_ do not
    reach the block already started (its flush reads the records
    asynchronously, by a scatter or by a full upload); the next block sees
    them."""
    n, ni = 3000, 32
    s = plummer_state(n, seed=77)
    load_j(g, s)

    def send(a, x):
        assert g.set_j_particle(a, int(s["idx"][a]), s["tj"][a], s["dtj"][a], s["m"][a],
                                s["k18"][a], s["j6"][a], s["a2"][a], s["v"][a], x) == 0

    for a in range(pre_dirty):            # dirty records for this block's flush
        send(a, s["x"][a])
    g.set_ti(0.0)
    xp, vp = oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], 0.0, 4)
    isel = np.arange(ni)
    g.calc_firsthalf(n, s["idx"][isel], xp[isel], vp[isel], 1e-8, np.zeros(ni))
    moved = np.arange(50, 250)
    for a in moved:
        send(a, s["x"][a] + 5.0)
    rc, acc, jerk, pot, nn = g.calc_lasthalf2(n, ni)
    assert rc == 0, g.last_error()
    ref = oracle.force(xp[isel], vp[isel], s["idx"][isel], xp, vp, s["m"], s["idx"], 1e-8)
    errs = compare(acc, jerk, pot, nn, ref, xp[isel], vp[isel], s["m"], xp, vp, 1e-8, isel)
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6, errs
    s2 = dict(s, x=s["x"].copy())
    s2["x"][moved] += 5.0
    errs = _block(g, s2, isel)
    assert errs["nn_mismatch"] == 0 and errs["acc_norm"] < 1e-6, errs
