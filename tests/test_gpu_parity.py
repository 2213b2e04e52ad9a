"""GPU parity: libg6.so (HIP, gfx950) through its C ABI against the FP64
CPU oracle on identical inputs.

Bars (see tests/gpu_common.py): predicted positions and velocities bitwise
equal; pot relative error <= 1e-6; acc and jerk normwise error (|delta| over
the sum of the magnitudes of the pair terms) <= 1e-6; per-particle acc
error |da|/|a| <= 1e-6 (Plummer inputs); nearest-neighbour index and
neighbour-list membership and order exactly equal.
"""
import numpy as np
import pytest

import oracle
from tests.gpu_common import (ACC_REL_TOL, ACC_TOL, JERK_TOL, POT_TOL, compare, load_j,
                              per_particle, plummer_state)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from starlab_amd.g6 import GRAPE6
    dev = GRAPE6()
    assert dev.open() == 0, dev.last_error()
    yield dev
    dev.close()


@pytest.fixture(params=[(0, 64), (0, 128), (0, 256), (1, 256)],
                ids=["smem64", "smem128", "smem256", "lds"])
def variant(g, request):
    """Run a test on the scalar-load kernel with each j-tile length and on the
    LDS-staged kernel."""
    v, tile = request.param
    g.option("force_variant", v)
    g.option("tile", tile)
    yield v
    g.option("force_variant", -1)
    g.option("tile", 0)


def _predict(s, ti, order=4):
    return oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], ti, order)


def _check(errs, label):
    print(label, errs, flush=True)
    assert errs["pot_rel"] <= POT_TOL, errs
    assert errs["acc_norm"] <= ACC_TOL, errs
    assert errs["jerk_norm"] <= JERK_TOL, errs
    assert errs["acc_rel_max"] <= ACC_REL_TOL, errs
    assert errs["nn_mismatch"] == 0, errs


def _abi_force_vs_oracle(g, s, ti, eps2, isel, order=4, excl=None, xi=None, vi=None):
    nj = len(s["m"])
    g.set_ti(ti)
    xp, vp = _predict(s, ti, order)
    idx = s["idx"]
    xi = xp[isel] if xi is None else xi
    vi = vp[isel] if vi is None else vi
    ex = idx[isel] if excl is None else np.asarray(excl, np.int32)
    acc, jerk, pot, nn = g.force(nj, ex, xi, vi, eps2)
    ref = oracle.force(xi, vi, ex, xp, vp, s["m"], idx, eps2)
    ex_addr = np.array([e if 0 <= e < nj else -1 for e in ex])
    return compare(acc, jerk, pot, nn, ref, xi, vi, s["m"], xp, vp, eps2, ex_addr)


def test_predictor_bitwise(g):
    s = plummer_state(1000, seed=3, k18=True)
    load_j(g, s)
    for order in (4, 5):
        g.option("predict_order", order)
        ti = 0.0
        g.set_ti(ti)
        idx = s["idx"][:4]
        g.force(1000, idx, s["x"][:4], s["v"][:4], 1e-8)
        xp, vp = g.read_predicted(1000)
        rx, rv = _predict(s, ti, order)
        np.testing.assert_array_equal(xp, rx)
        np.testing.assert_array_equal(vp, rv)
    g.option("predict_order", 4)


def test_plummer_1024_per_particle_abi(g, variant):
    """Config-1 shape: N=1024 Plummer, eps = 0, loaded one particle at a time
    through g6_set_j_particle_ exactly as kira does."""
    s = plummer_state(1024, seed=42)
    load_j(g, s, bulk=False)
    errs = _abi_force_vs_oracle(g, s, 0.0, 0.0, np.arange(1024))
    _check(errs, "plummer1024 eps=0")


def test_chunking_and_ragged_sizes(g, variant):
    old = g.npipes()
    s = plummer_state(1000, seed=5)      # nj not a multiple of the 256-j tile
    load_j(g, s)
    try:
        for npipes, isel in ((100, np.arange(1000)), (7, np.arange(1, 20)), (2048, np.array([999])),
                             (2048, np.arange(0, 1000, 3))):
            g.option("npipes", npipes)
            errs = _abi_force_vs_oracle(g, s, 0.0, 1e-6, isel)
            _check(errs, f"npipes={npipes} ni={len(isel)}")
    finally:
        g.option("npipes", old)


def test_tiny_systems(g, variant):
    for n in (1, 2, 3, 65):
        s = plummer_state(max(n, 2), seed=11)
        for k in s:
            s[k] = s[k][:n]
        load_j(g, s)
        acc, jerk, pot, nn = g.force(n, s["idx"], s["x"], s["v"], 0.0)
        xp, vp = _predict(s, 0.0)
        racc, rjerk, rpot, rnn, _ = oracle.force(s["x"], s["v"], s["idx"], xp, vp, s["m"],
                                                s["idx"], 0.0)
        if n == 1:
            assert np.all(acc == 0) and np.all(jerk == 0) and pot[0] == 0 and nn[0] == 0
            continue
        np.testing.assert_allclose(pot, rpot, rtol=1e-6)
        np.testing.assert_allclose(acc, racc, rtol=1e-5, atol=1e-6 * np.abs(racc).max())
        np.testing.assert_array_equal(nn, rnn)


def test_predictor_time_spread_and_order5(g, variant):
    s = plummer_state(2000, seed=7, k18=True)
    load_j(g, s)
    ti = 0.0
    for order in (4, 5):
        g.option("predict_order", order)
        errs = _abi_force_vs_oracle(g, s, ti, 1e-8, np.arange(0, 2000, 2), order=order)
        _check(errs, f"order{order}")
    g.option("predict_order", 4)


def test_exclusion_by_parent_index(g, variant):
    """Perturbation call (hdyn_grape6.C:3415-3558): two components share the
    parent's index; that j is excluded for both, positions are off-grid."""
    s = plummer_state(600, seed=9)
    load_j(g, s)
    xp, vp = _predict(s, 0.0)
    xi = xp[[10, 10, 20]] + np.array([[1e-4, 0, 0], [-1e-4, 0, 0], [0, 0, 0]])
    vi = vp[[10, 10, 20]]
    errs = _abi_force_vs_oracle(g, s, 0.0, 0.0, None, excl=[10, 10, 20], xi=xi, vi=vi)
    _check(errs, "parent-index exclusion")
    # an index that is not in j-memory excludes nothing
    xi2 = xp[[5]] + 1e-3
    errs = _abi_force_vs_oracle(g, s, 0.0, 0.0, None, excl=[123456], xi=xi2, vi=vp[[5]])
    _check(errs, "no exclusion")


def test_close_pair_term_in_fp64(g, variant):
    """A pair 1e-4 apart inside a Plummer sphere: its mutual acc (~1e5) and
    jerk outweigh the rest of the sum 1e5 to 1.  The reduction recomputes the
    nearest-neighbour tile in FP64, so both members agree with the FP64
    oracle to ~1e-11 relative; FP32 pair arithmetic alone leaves ~1e-7
    (rsqrt and r^-3 rounding of the dominant term).  Both members, every
    kernel variant and tile length, the pair inside one tile and split over
    two."""
    n = 1024
    for a, b in ((300, 301), (300, 700)):
        s = plummer_state(n, seed=5)
        s["x"][b] = s["x"][a] + np.array([1e-4, 0.6e-4, -0.3e-4])
        s["v"][b] = s["v"][a] + np.array([0.0, 0.01, 0.02])
        s["tj"][[a, b]] = 0.0   # predicted to ti = 0 unchanged: the separation stays
        load_j(g, s)
        isel = np.arange(n)
        errs = _abi_force_vs_oracle(g, s, 0.0, 0.0, isel)
        _check(errs, f"close pair {a},{b}")
        xp, vp = _predict(s, 0.0)
        acc, jerk, pot, nn = g.force(n, s["idx"][[a, b]], xp[[a, b]], vp[[a, b]], 0.0)
        racc, rjerk, rpot, rnn, _ = oracle.force(xp[[a, b]], vp[[a, b]], s["idx"][[a, b]], xp,
                                                 vp, s["m"], s["idx"], 0.0)
        ra, rj, rp = per_particle(acc, jerk, pot, racc, rjerk, rpot)
        print(f"close pair {a},{b}: acc {ra} jerk {rj} pot {rp}", flush=True)
        assert list(nn) == [b, a]
        assert ra.max() <= 1e-9 and rj.max() <= 1e-9, (ra, rj)
        assert rp.max() <= POT_TOL


def test_nn_ties_on_lattice(g, variant):
    gr = np.arange(6, dtype=float)
    x = np.array([[a, b, c] for a in gr for b in gr for c in gr]) * 0.125
    n = len(x)
    s = dict(m=np.full(n, 1.0 / n), x=x, v=np.zeros_like(x), a2=np.zeros_like(x),
             j6=np.zeros_like(x), k18=np.zeros_like(x), tj=np.zeros(n), dtj=np.ones(n),
             idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    g.set_ti(0.0)
    acc, jerk, pot, nn = g.force(n, s["idx"], x, s["v"], 0.0)
    _, _, _, rnn, _ = oracle.force(x, s["v"], s["idx"], x, s["v"], s["m"], s["idx"], 0.0)
    np.testing.assert_array_equal(nn, rnn)
    assert g.refinements() > 0        # exact ties must go through the FP64 rescan


def test_nn_ties_across_splits(g, variant):
    """A 16^3 lattice in random address order: every point's nearest
    neighbours (up to six at one exact distance) sit in different j-splits.
    The FP64 rescan visits only the splits whose own FP32 best lies within
    the band of the best, and must still return the lowest address among the
    tied neighbours, as the oracle does."""
    gr = np.arange(16, dtype=float)
    x = np.array([[a, b, c] for a in gr for b in gr for c in gr]) * 0.125
    x = x[np.random.default_rng(3).permutation(len(x))]
    n = len(x)
    s = dict(m=np.full(n, 1.0 / n), x=x, v=np.zeros_like(x), a2=np.zeros_like(x),
             j6=np.zeros_like(x), k18=np.zeros_like(x), tj=np.zeros(n), dtj=np.ones(n),
             idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    g.set_ti(0.0)
    _, _, _, nn = g.force(n, s["idx"], x, s["v"], 0.0)
    _, nsplit, jchunk = g.geometry()       # of the last call
    print("geometry: nsplit", nsplit, "jchunk", jchunk, "refined", g.refinements(), flush=True)
    _, _, _, rnn, _ = oracle.force(x, s["v"], s["idx"], x, s["v"], s["m"], s["idx"], 0.0)
    np.testing.assert_array_equal(nn, rnn)
    assert nsplit > 1 and g.refinements() > 0


def test_neighbour_lists_exact(g):
    s = plummer_state(3000, seed=13)
    load_j(g, s)
    ti = 0.0
    g.set_ti(ti)
    xp, vp = _predict(s, ti)
    isel = np.arange(0, 3000, 5)[:512]
    rng = np.random.default_rng(1)
    h2 = rng.uniform(0.0, 0.02, len(isel))
    h2[::4] = 0.0
    g.calc_firsthalf(3000, s["idx"][isel], xp[isel], vp[isel], 1e-8, h2)
    rc, *_ = g.calc_lasthalf2(3000, len(isel))
    assert rc == 0
    assert g.read_neighbour_list() == 0
    counts, lists = oracle.neighbours(xp[isel], h2, s["idx"][isel], 1, xp, s["idx"], 4096)
    for k in range(len(isel)):
        rc, lst = g.get_neighbour_list(k, 4096)
        assert rc == 0
        assert len(lst) == counts[k]
        np.testing.assert_array_equal(lst, lists[k, :counts[k]])
        if h2[k] > 0:
            assert s["idx"][isel][k] in lst      # n_pipes > 48: self is listed


@pytest.mark.parametrize("base", [0.25, 10.0, 1000.0])
def test_neighbour_lists_exact_on_the_boundary(g, base):
    """A 2^-10 lattice (a third of its points jittered by 1e-3 of a step)
    at coordinates around `base`, where FP32 resolves 3e-8 (0.25) to 6e-5
    (1000) of a coordinate: half the radii are exactly one step (lattice
    neighbours at d2 == h2, which the strict test leaves out), half random.
    Both list sweeps compare FP32 r^2 first and leave a band around h2 to
    the FP64 test; the lists must be the oracle's exactly, including where
    FP32 cannot tell any pair apart (base 1000)."""
    a = 2.0 ** -10
    gr = np.arange(12) * a
    lat = np.array([[p, q, r] for p in gr for q in gr for r in gr])
    rng = np.random.default_rng(7)
    jit = rng.normal(0.0, a * 1e-3, lat.shape) * (np.arange(len(lat)) % 3 == 0)[:, None]
    x = base + lat + jit
    n = len(x)
    z = np.zeros_like(x)
    s = dict(m=np.full(n, 1.0 / n), x=x, v=z, a2=z, j6=z, k18=z, tj=np.zeros(n),
             dtj=np.ones(n), idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    g.set_ti(0.0)
    isel = np.arange(1, n, 3)[:512]             # not jittered
    h2 = rng.uniform(0.5, 3.0, len(isel)) * a * a
    h2[::2] = a * a
    g.calc_firsthalf(n, s["idx"][isel], x[isel], z[isel], 0.0, h2)
    rc, *_ = g.calc_lasthalf2(n, len(isel))
    assert rc == 0
    assert g.read_neighbour_list() == 0
    counts, lists = oracle.neighbours(x[isel], h2, s["idx"][isel], 1, x, s["idx"], 4096)
    assert counts.min() >= 1 and counts.max() > 10
    for k in range(len(isel)):
        rc, lst = g.get_neighbour_list(k, 4096)
        assert rc == 0
        assert len(lst) == counts[k], (k, len(lst), counts[k])
        np.testing.assert_array_equal(lst, lists[k, :counts[k]])


def test_neighbour_list_overflow(g):
    s = plummer_state(2000, seed=17)
    load_j(g, s)
    g.set_ti(0.0)
    xp, vp = _predict(s, 0.0)
    isel = np.array([0, 1, 2])
    h2 = np.array([100.0, 0.01, 0.0])              # first one holds every j
    g.calc_firsthalf(2000, s["idx"][isel], xp[isel], vp[isel], 0.0, h2)
    assert g.calc_lasthalf2(2000, 3)[0] == 0
    assert g.read_neighbour_list() == 0            # 2000 <= nb_cap
    rc, lst = g.get_neighbour_list(0, 256)          # caller buffer too small
    assert rc == 1 and len(lst) == 256
    np.testing.assert_array_equal(lst, np.arange(256))
    rc, lst = g.get_neighbour_list(2, 256)
    assert rc == 0 and len(lst) == 0
    old = g.option("nb_cap")
    g.option("nb_cap", 100)                         # device storage too small
    try:
        g.calc_firsthalf(2000, s["idx"][isel], xp[isel], vp[isel], 0.0, h2)
        assert g.calc_lasthalf2(2000, 3)[0] == 0
        assert g.read_neighbour_list() > 0
        rc, lst = g.get_neighbour_list(0, 4096)
        assert rc == 1
    finally:
        g.option("nb_cap", old)


def test_lazy_j_updates_and_reload(g):
    s = plummer_state(1500, seed=19)
    load_j(g, s)
    isel = np.arange(0, 1500, 7)
    errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
    _check(errs, "before update")
    rng = np.random.default_rng(2)
    moved = rng.choice(1500, 40, replace=False)
    for a in moved:
        s["x"][a] += rng.normal(scale=0.01, size=3)
        s["tj"][a] = -1e-4
        g.set_j_particle(int(a), int(a), s["tj"][a], 1.0, s["m"][a], s["k18"][a], s["j6"][a],
                         s["a2"][a], s["v"][a], s["x"][a])
    errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
    _check(errs, "after scatter update")
    # close / reopen (GRAPE release and re-attach): kira reloads j
    assert g.close() == 0
    assert g.open() == 0
    load_j(g, s)
    errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
    _check(errs, "after reopen")
    assert g.reset() == 0
    load_j(g, s)
    errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
    _check(errs, "after reset")


def test_lasthalf_without_firsthalf_is_an_error(g):
    rc, *_ = g.calc_lasthalf2(10, 1)
    assert rc != 0


def test_self_sweep_matches_abi_and_is_deterministic(g, variant):
    s = plummer_state(4096, seed=23)
    load_j(g, s)
    ti = 0.0
    assert g.self_sweep(4096, 0, 4096, ti, 1e-8) == 0
    a1, j1, p1, n1 = g.read_results(4096)
    assert g.self_sweep(4096, 0, 4096, ti, 1e-8) == 0
    a2, j2, p2, n2 = g.read_results(4096)
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(j1, j2)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(n1, n2)
    xp, vp = _predict(s, ti)
    ref = oracle.force(xp, vp, s["idx"], xp, vp, s["m"], s["idx"], 1e-8)
    errs = compare(a1, j1, p1, n1, ref, xp, vp, s["m"], xp, vp, 1e-8, s["idx"])
    _check(errs, "self sweep 4096")


def test_full_size_65536_sample(g):
    """The force-only benchmark workload: Plummer N=65536, eps2=1e-8, t_j
    spread over [0, 2^-10].  Checked on a 256-i sample against the oracle,
    plus size-independent properties of the whole sweep."""
    from starlab_amd.ics import makeplummer
    n = 65536
    m, x, v = makeplummer(n, 65536, potential_fn=lambda m, x: -0.25 * 2.0)
    rng = np.random.default_rng(0)
    s = dict(m=m, x=x, v=v, a2=rng.normal(size=(n, 3)) * 0.5, j6=rng.normal(size=(n, 3)),
             k18=np.zeros((n, 3)), tj=-rng.uniform(0, 2.0 ** -10, n), dtj=np.ones(n),
             idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    assert g.self_sweep(n, 0, n, 0.0, 1e-8) == 0
    acc, jerk, pot, nn = g.read_results(n)
    xp, vp = _predict(s, 0.0)
    sel = rng.choice(n, 256, replace=False)
    ref = oracle.force(xp[sel], vp[sel], s["idx"][sel], xp, vp, m, s["idx"], 1e-8)
    errs = compare(acc[sel], jerk[sel], pot[sel], nn[sel], ref, xp[sel], vp[sel], m, xp, vp,
                   1e-8, s["idx"][sel])
    _check(errs, "65536 sample")
    # momentum: sum m a = 0 up to the FP32 pair rounding
    ma = (m[:, None] * acc).sum(0)
    assert np.abs(ma).max() < 1e-6 * (m[:, None] * np.abs(acc)).sum(0).max()
    # potential energy agrees with the oracle's pair sum
    ep, _ = oracle.energies(m, xp, vp, 1e-8)
    assert abs(0.5 * (m * pot).sum() - ep) < 1e-7 * abs(ep)
    assert np.all(pot < 0) and np.all((nn >= 0) & (nn < n))
    _whole_sweep_properties(m, xp, acc, jerk, nn)


def _whole_sweep_properties(m, xp, acc, jerk, nn):
    """Size-independent checks of a full self-sweep: Newton's third law for
    acc and jerk (sum m a = sum m j = 0 up to the FP32 pair rounding), and
    every particle's nearest neighbour at the exact FP64 nearest distance
    (a k-d tree over all positions; distances are compared, so an exact tie
    may resolve to either index)."""
    from scipy.spatial import cKDTree
    mj = (m[:, None] * jerk).sum(0)
    assert np.abs(mj).max() < 1e-6 * (m[:, None] * np.abs(jerk)).sum(0).max()
    d, k = cKDTree(xp).query(xp, k=2)
    d_nn = np.sqrt(((xp[nn] - xp) ** 2).sum(1))
    assert np.all(nn != np.arange(len(m)))
    bad = np.flatnonzero(np.abs(d_nn - d[:, 1]) > 1e-12 * d[:, 1])
    assert len(bad) == 0, f"{len(bad)} nearest neighbours off, e.g. i={bad[:5]}"


def test_full_size_262144_sweep_properties(g):
    """Config 3's j-memory size: Plummer N = 262144 swept in full, checked on
    a 128-i sample against the oracle and on every particle through the
    size-independent properties."""
    from starlab_amd.ics import makeplummer
    n = 262144
    m, x, v = makeplummer(n, 262144, potential_fn=lambda m, x: -0.25 * 2.0)
    rng = np.random.default_rng(1)
    s = dict(m=m, x=x, v=v, a2=rng.normal(size=(n, 3)) * 0.5, j6=rng.normal(size=(n, 3)),
             k18=np.zeros((n, 3)), tj=-rng.uniform(0, 2.0 ** -10, n), dtj=np.ones(n),
             idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    assert g.self_sweep(n, 0, n, 0.0, 0.0) == 0
    acc, jerk, pot, nn = g.read_results(n)
    xp, vp = _predict(s, 0.0)
    sel = rng.choice(n, 128, replace=False)
    ref = oracle.force(xp[sel], vp[sel], s["idx"][sel], xp, vp, m, s["idx"], 0.0)
    errs = compare(acc[sel], jerk[sel], pot[sel], nn[sel], ref, xp[sel], vp[sel], m, xp, vp,
                   0.0, s["idx"][sel])
    _check(errs, "262144 sample")
    ma = (m[:, None] * acc).sum(0)
    assert np.abs(ma).max() < 1e-6 * (m[:, None] * np.abs(acc)).sum(0).max()
    _whole_sweep_properties(m, xp, acc, jerk, nn)


@pytest.mark.parametrize("nsplit", [0, 7, 300])
def test_small_blocks_many_splits(g, nsplit):
    """kira's small blocks: a few i against many j, spread over many j-splits
    (more splits than lanes of the reduction wave, a split count that does
    not divide the j range, and the automatic choice)."""
    s = plummer_state(20000, seed=29)
    load_j(g, s)
    g.option("nsplit", nsplit)
    g.option("small_fp64", 0)   # the FP32 kernels' shapes (blocks of <= 16 i default to FP64)
    try:
        rng = np.random.default_rng(5)
        for ni in (1, 3, 64, 65):
            isel = rng.choice(20000, ni, replace=False)
            errs = _abi_force_vs_oracle(g, s, 0.0, 0.0, isel)
            _check(errs, f"ni={ni} nsplit={nsplit} tile={g.option('last_tile')}")
    finally:
        g.option("nsplit", 0)
        g.option("small_fp64", 1)


@pytest.mark.parametrize("tile", [16, 32])
def test_short_tiles(g, tile):
    """Short j-tiles (the small-block shape's 16 and 32 j per FP32 tile) on a
    j count that is not a multiple of the tile, for blocks of one i-group and
    of several (self-exclusion, exact nearest neighbours, softened and not)."""
    s = plummer_state(9001, seed=43)
    load_j(g, s)
    g.option("force_variant", 0)
    g.option("tile", tile)
    g.option("small_fp64", 0)
    try:
        rng = np.random.default_rng(tile)
        for ni, eps2 in ((1, 0.0), (40, 1e-8), (300, 0.0)):
            isel = rng.choice(9001, ni, replace=False)
            errs = _abi_force_vs_oracle(g, s, 0.0, eps2, isel)
            assert g.option("last_tile") == tile
            _check(errs, f"tile={tile} ni={ni}")
    finally:
        g.option("tile", 0)
        g.option("force_variant", -1)
        g.option("small_fp64", 1)


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_reduce_lanes_per_i(g, lanes):
    """Every lanes-per-i shape of the split reduction, on a split count (47)
    that is not a multiple of any of them and an i count that leaves partial
    waves."""
    s = plummer_state(6000, seed=31)
    load_j(g, s)
    g.option("nsplit", 47)
    g.option("reduce_lanes", lanes)
    try:
        isel = np.random.default_rng(lanes).choice(6000, 300, replace=False)
        errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
        _check(errs, f"lanes={lanes}")
    finally:
        g.option("nsplit", 0)
        g.option("reduce_lanes", 0)


@pytest.mark.parametrize("wpg", [1, 2, 4, 8, 16])
def test_waves_per_workgroup_fold(g, wpg):
    """Every waves-per-workgroup count of the scalar-load kernel (the waves of
    a workgroup fold their j-splits in LDS before the reduction), on split
    counts that leave a workgroup partly empty (47) and that divide evenly
    (64), with index exclusion and exact nearest neighbours."""
    s = plummer_state(6000, seed=41)
    load_j(g, s)
    g.option("force_variant", 0)
    g.option("wpg", wpg)
    try:
        for nsplit in (47, 64):
            g.option("nsplit", nsplit)
            isel = np.random.default_rng(wpg + nsplit).choice(6000, 200, replace=False)
            errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
            assert g.option("last_wpg") == wpg
            _check(errs, f"wpg={wpg} nsplit={nsplit}")
    finally:
        g.option("nsplit", 0)
        g.option("wpg", 0)
        g.option("force_variant", -1)


def test_fold_is_deterministic_and_matches_unfolded(g):
    """The same block twice gives identical bits for every wpg, and folding
    changes results only by FP64 reassociation of the split sums."""
    s = plummer_state(8000, seed=43)
    load_j(g, s)
    xp, vp = _predict(s, 0.0)
    isel = np.arange(0, 8000, 13)
    g.option("force_variant", 0)
    g.option("nsplit", 64)
    outs = {}
    try:
        for wpg in (1, 4, 16):
            g.option("wpg", wpg)
            a = g.force(8000, s["idx"][isel], xp[isel], vp[isel], 1e-8)
            b = g.force(8000, s["idx"][isel], xp[isel], vp[isel], 1e-8)
            for u, w in zip(a, b):
                np.testing.assert_array_equal(u, w)
            outs[wpg] = a
    finally:
        g.option("nsplit", 0)
        g.option("wpg", 0)
        g.option("force_variant", -1)
    for wpg in (4, 16):
        np.testing.assert_allclose(outs[wpg][0], outs[1][0], rtol=0, atol=1e-12 * np.abs(outs[1][0]).max())
        np.testing.assert_allclose(outs[wpg][2], outs[1][2], rtol=1e-13)
        np.testing.assert_array_equal(outs[wpg][3], outs[1][3])


def test_autotune_table_and_parity(g, tmp_path):
    """The autotuner fills one shape per i-block bucket, forces computed with
    tuned shapes keep parity, and the table survives a save/load round
    trip."""
    s = plummer_state(20000, seed=37)
    load_j(g, s)
    try:
        n = g.autotune(20000, 700, 1e-8)
        tab = g.tuning()
        assert n == len(tab) == 11                      # buckets 1, 2, 4, ..., 1024 (ni <= 700)
        for b, (v, tile, waves, lanes) in tab.items():
            assert v in (0, 1) and tile in (16, 32, 64, 128, 256) and waves > 0
            assert lanes in (0, 4, 8, 16, 32, 64)         # 0: the split-count rule won
        rng = np.random.default_rng(9)
        for ni in (1, 5, 64, 300, 700):
            isel = rng.choice(20000, ni, replace=False)
            errs = _abi_force_vs_oracle(g, s, 0.0, 1e-8, isel)
            _check(errs, f"tuned ni={ni} shape={tab[max(0, int(np.ceil(np.log2(ni))))]}")
        path = tmp_path / "shapes.txt"
        assert g.save_tuning(path) == 0
        g.clear_tuning()
        assert g.tuning() == {}
        assert g.load_tuning(path) == len(tab)
        assert g.tuning() == tab
    finally:
        g.clear_tuning()


@pytest.mark.parametrize("nj", [1000, 65536])
@pytest.mark.parametrize("eps2", [1e-8, 0.0])
def test_fused_small_blocks_equal_the_predictor_path(g, nj, eps2):
    """ABI blocks of up to four i-groups predict j inside the force kernel
    (257 i is the first size that does not).  Their
    acc, jerk, pot, nearest neighbours, neighbour lists and the FP64
    prediction read back are bitwise those of the predictor + JPred path
    (same pair order, same partial sums), and they pass the oracle checks."""
    s = plummer_state(nj, seed=71)
    load_j(g, s)
    rng = np.random.default_rng(nj)
    ti = 0.0
    xp, vp = _predict(s, ti)
    g.option("small_fp64", 0)   # blocks of up to 16 i would take the FP64 kernel
    try:
        for ni in (1, 13, 64, 65, 200, 256, 257):
            isel = np.sort(rng.choice(nj, ni, replace=False))
            h2 = np.full(ni, 0.05 ** 2)
            out = {}
            for fused in (1, 0):
                g.option("fused_small", fused)
                g.set_ti(ti)
                g.calc_firsthalf(nj, s["idx"][isel], xp[isel], vp[isel], eps2, h2)
                rc, acc, jerk, pot, nn = g.calc_lasthalf2(nj, ni)
                assert rc == 0, g.last_error()
                assert g.read_neighbour_list() == 0
                lists = [g.get_neighbour_list(k, 8192)[1] for k in range(ni)]
                xr, vr = g.read_predicted(nj)
                out[fused] = (acc, jerk, pot, nn, xr, vr, lists)
            for u, w, name in zip(out[1][:6], out[0][:6], ("acc", "jerk", "pot", "nn", "xp", "vp")):
                np.testing.assert_array_equal(u, w, err_msg=f"ni={ni} {name}")
            for k, (u, w) in enumerate(zip(out[1][6], out[0][6])):
                np.testing.assert_array_equal(u, w, err_msg=f"ni={ni} list {k}")
            np.testing.assert_array_equal(out[1][4], xp)
            ref = oracle.force(xp[isel], vp[isel], s["idx"][isel], xp, vp, s["m"], s["idx"], eps2)
            acc, jerk, pot, nn = out[1][:4]
            _check(compare(acc, jerk, pot, nn, ref, xp[isel], vp[isel], s["m"], xp, vp, eps2,
                           isel), f"fused nj={nj} ni={ni}")
    finally:
        g.option("fused_small", 1)
        g.option("small_fp64", 1)


def test_large_abi_block_equals_self_sweep(g):
    """One ABI block of 20000 i (npipes raised; the scatter launch reads the
    1.6 MB i-block from pinned memory) gives bitwise the results of the
    device-resident sweep over the same particles: the host and device
    predictions are bitwise equal, so both build the same double-single
    i-block and run the same force geometry."""
    n = 20000
    s = plummer_state(n, seed=73)
    load_j(g, s)
    xp, vp = _predict(s, 0.0)
    old = g.npipes()
    try:
        g.option("npipes", n)
        g.set_ti(0.0)
        acc, jerk, pot, nn = g.force(n, s["idx"], xp, vp, 1e-8)
        assert g.self_sweep(n, 0, n, 0.0, 1e-8) == 0
        a2, j2, p2, n2 = g.read_results(n)
    finally:
        g.option("npipes", old)
    for u, w, name in ((acc, a2, "acc"), (jerk, j2, "jerk"), (pot, p2, "pot"), (nn, n2, "nn")):
        np.testing.assert_array_equal(u, w, err_msg=name)
    sel = np.random.default_rng(5).choice(n, 128, replace=False)
    ref = oracle.force(xp[sel], vp[sel], s["idx"][sel], xp, vp, s["m"], s["idx"], 1e-8)
    _check(compare(acc[sel], jerk[sel], pot[sel], nn[sel], ref, xp[sel], vp[sel], s["m"], xp, vp,
                   1e-8, sel), "ABI block of 20000")
