"""Small ABI blocks -- at most 12 i with one softening, kira's commonest
calls -- run as one launch (k_small): FP64 pair arithmetic with the
j-particles spread over the lanes, the split reduction inside the kernel,
and pending j updates carried in the launch's arguments.

Against the FP64 oracle on identical inputs: per-particle acc and pot within
1e-12 relative, jerk within 1e-13 normwise (the FP64 pair law in another
summation order), nearest neighbours exact, ties included.  Neighbour lists
and the FP64 prediction read back equal the FP32 path's bit for bit.  j
updates that the launch carried reach HBM for later blocks of any size, and
repeated calls are bitwise equal."""
import numpy as np
import pytest

import oracle
from tests.gpu_common import compare, load_j, plummer_state

pytestmark = pytest.mark.gpu

ACC64 = 1e-12
JERK64 = 1e-13
POT64 = 1e-12


@pytest.fixture(scope="module")
def g():
    from starlab_amd.g6 import GRAPE6
    dev = GRAPE6()
    assert dev.open() == 0, dev.last_error()
    assert dev.option("small_fp64") == 1
    yield dev
    dev.close()


def sphere_state(n, seed):
    """n equal masses uniform in the unit sphere, random velocities, and
    random predictor inputs (for sizes where a Plummer draw is slow)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    x *= (rng.uniform(size=n) ** (1 / 3) / np.linalg.norm(x, axis=1))[:, None]
    return dict(m=np.full(n, 1.0 / n), x=x, v=rng.normal(scale=0.5, size=(n, 3)),
                a2=rng.normal(size=(n, 3)), j6=rng.normal(size=(n, 3)), k18=np.zeros((n, 3)),
                tj=-rng.uniform(0.0, 2.0 ** -10, n), dtj=np.full(n, 2.0 ** -10),
                idx=np.arange(n, dtype=np.int32))


def _predict(s, ti):
    return oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], ti, 4)


def _check64(acc, jerk, pot, nn, s, xp, vp, xi, vi, excl, eps2, label):
    ref = oracle.force(xi, vi, excl, xp, vp, s["m"], s["idx"], eps2)
    ex_addr = np.array([e if 0 <= e < len(s["m"]) else -1 for e in excl])
    errs = compare(acc, jerk, pot, nn, ref, xi, vi, s["m"], xp, vp, eps2, ex_addr)
    print(label, {k: errs[k] for k in ("acc_rel_max", "jerk_norm", "pot_rel", "nn_mismatch")},
          flush=True)
    assert errs["acc_rel_max"] <= ACC64, (label, errs)
    assert errs["jerk_norm"] <= JERK64, (label, errs)
    assert errs["pot_rel"] <= POT64, (label, errs)
    assert errs["nn_mismatch"] == 0, (label, errs)


@pytest.mark.parametrize("nj", [1000, 65536, 262144, 300000, 600000])
def test_small_blocks_match_fp64_oracle(g, nj):
    """About nj/512 workgroups of 256 threads, at most 256, 64 j per wave and
    chunk: 2, 2, 4, 5 and 10 chunks per wave (the next chunk's records load
    while one is summed); i counts 1, 2, 3, 5, 8, 9, 12 cover every lane
    grouping (1 to 16 i-slots), and 13 takes the FP32 path."""
    wgs = min(256, -(-nj // 512))
    chunks = -(-nj // (wgs * 256))
    s = plummer_state(nj, seed=91) if nj <= 65536 else sphere_state(nj, 91)
    load_j(g, s)
    rng = np.random.default_rng(nj)
    ti = 2.0 ** -12
    g.set_ti(ti)
    xp, vp = _predict(s, ti)
    for ni in (1, 2, 3, 5, 8, 9, 12):
        isel = np.sort(rng.choice(nj, ni, replace=False))
        for eps2 in (1e-8, 0.0):
            acc, jerk, pot, nn = g.force(nj, s["idx"][isel], xp[isel], vp[isel], eps2)
            assert g.L.g6x_last_force_geometry(None, None, None) == 2   # k_small ran
            assert g.geometry() == (wgs, wgs, 256 * chunks)
            _check64(acc, jerk, pot, nn, s, xp, vp, xp[isel], vp[isel], s["idx"][isel], eps2,
                     f"nj={nj} ni={ni} eps2={eps2}")
    # an i that is not in j-memory (nothing excluded), away from every j
    xi = np.array([[2.5, -1.0, 0.5]])
    vi = np.array([[0.1, 0.2, -0.3]])
    excl = np.array([10 ** 7], np.int32)
    acc, jerk, pot, nn = g.force(nj, excl, xi, vi, 1e-8)
    _check64(acc, jerk, pot, nn, s, xp, vp, xi, vi, excl, 1e-8, f"nj={nj} outside i")
    isel = np.sort(rng.choice(nj, 13, replace=False))
    g.force(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8)
    assert g.L.g6x_last_force_geometry(None, None, None) != 2   # 13 i: the FP32 path


def test_small_block_nn_ties_exact(g):
    gr = np.arange(6, dtype=float)
    x = np.array([[a, b, c] for a in gr for b in gr for c in gr]) * 0.125
    n = len(x)
    z = np.zeros_like(x)
    s = dict(m=np.full(n, 1.0 / n), x=x, v=z, a2=z, j6=z, k18=z, tj=np.zeros(n),
             dtj=np.ones(n), idx=np.arange(n, dtype=np.int32))
    load_j(g, s)
    g.set_ti(0.0)
    for first in range(0, n, 8):
        sel = np.arange(first, min(n, first + 8))
        acc, jerk, pot, nn = g.force(n, s["idx"][sel], x[sel], z[sel], 0.0)
        _, _, _, rnn, _ = oracle.force(x[sel], z[sel], s["idx"][sel], x, z, s["m"], s["idx"],
                                       0.0)
        np.testing.assert_array_equal(nn, rnn)
        assert g.refinements() == 0   # FP64 distances: no rescan


def test_small_block_lists_and_prediction_equal_fp32_path(g):
    nj = 65536
    s = plummer_state(nj, seed=93)
    load_j(g, s)
    rng = np.random.default_rng(3)
    ti = 0.0
    xp, vp = _predict(s, ti)
    listed = 0
    try:
        for ni in (1, 5, 8):
            isel = np.sort(rng.choice(nj, ni, replace=False))
            h2 = np.full(ni, 0.04 ** 2)
            out = {}
            for small in (1, 0):
                g.option("small_fp64", small)
                g.set_ti(ti)
                g.calc_firsthalf(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8, h2)
                rc, acc, jerk, pot, nn = g.calc_lasthalf2(nj, ni)
                assert rc == 0, g.last_error()
                assert g.read_neighbour_list() == 0
                lists = [g.get_neighbour_list(k, 8192)[1] for k in range(ni)]
                xr, vr = g.read_predicted(nj)
                out[small] = (acc, jerk, pot, nn, xr, vr, lists)
            np.testing.assert_array_equal(out[1][3], out[0][3], err_msg=f"ni={ni} nn")
            np.testing.assert_array_equal(out[1][4], out[0][4], err_msg=f"ni={ni} xp")
            np.testing.assert_array_equal(out[1][5], out[0][5], err_msg=f"ni={ni} vp")
            np.testing.assert_array_equal(out[1][4], xp)
            for k, (u, w) in enumerate(zip(out[1][6], out[0][6])):
                np.testing.assert_array_equal(u, w, err_msg=f"ni={ni} list {k}")
            listed += sum(len(u) for u in out[1][6]) - ni   # beyond each i itself
            _check64(*out[1][:4], s, xp, vp, xp[isel], vp[isel], s["idx"][isel], 1e-8,
                     f"lists ni={ni}")
        assert listed > 0   # the lists are not trivial
    finally:
        g.option("small_fp64", 1)


def test_small_block_carries_j_updates(g):
    """kira re-sends the nodes of the block it just advanced before the next
    block.  Up to 16 updates below nj travel in the small block's launch
    arguments; more, or an address at or beyond nj, go up by the scatter
    launch.  The small blocks hold 4 i.  Every later calculation -- small or large block, read-back --
    sees the updated records."""
    nj = 20000
    s = plummer_state(nj, seed=95)
    load_j(g, s)
    rng = np.random.default_rng(5)
    ti = 2.0 ** -11
    for step, k in enumerate((1, 5, 16, 17, 3)):
        addr = np.sort(rng.choice(nj - 1, k, replace=False))
        if step == 4:
            addr[-1] = nj - 1            # called with nj - 1 below: beyond the block's nj
        for a in addr:
            s["x"][a] += rng.normal(scale=1e-3, size=3)
            s["v"][a] += rng.normal(scale=1e-3, size=3)
            s["tj"][a] = ti - 2.0 ** -12
            assert g.set_j_particle(int(a), int(s["idx"][a]), s["tj"][a], s["dtj"][a],
                                    s["m"][a], s["k18"][a], s["j6"][a], s["a2"][a], s["v"][a],
                                    s["x"][a]) == 0
        n_call = nj - 1 if step == 4 else nj
        g.set_ti(ti)
        xp, vp = _predict(s, ti)
        sub = {k: v[:n_call] for k, v in s.items()}
        isel = np.sort(rng.choice(n_call, 4, replace=False))
        isel[0] = addr[0]                # an updated node is itself an i
        acc, jerk, pot, nn = g.force(n_call, s["idx"][isel], xp[isel], vp[isel], 1e-8)
        _check64(acc, jerk, pot, nn, sub, xp[:n_call], vp[:n_call], xp[isel], vp[isel],
                 s["idx"][isel], 1e-8, f"updates={k} small")
        # a large block (FP32 path) reads the records from HBM
        big = np.sort(rng.choice(nj, 300, replace=False))
        big[:len(addr)] = addr[:min(len(addr), 300)]
        big = np.unique(big)
        acc, jerk, pot, nn = g.force(nj, s["idx"][big], xp[big], vp[big], 1e-8)
        ref = oracle.force(xp[big], vp[big], s["idx"][big], xp, vp, s["m"], s["idx"], 1e-8)
        errs = compare(acc, jerk, pot, nn, ref, xp[big], vp[big], s["m"], xp, vp, 1e-8, big)
        assert errs["acc_norm"] <= 1e-6 and errs["pot_rel"] <= 1e-6, errs
        assert errs["nn_mismatch"] == 0, errs
        xr, vr = g.read_predicted(nj)
        np.testing.assert_array_equal(xr, xp)
        np.testing.assert_array_equal(vr, vp)
    # a re-open reloads every j and starts the small-block counter afresh
    assert g.close() == 0 and g.open() == 0
    g.set_ti(ti)
    isel = np.array([0, 7, 19999])
    acc, jerk, pot, nn = g.force(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8)
    _check64(acc, jerk, pot, nn, s, xp, vp, xp[isel], vp[isel], s["idx"][isel], 1e-8,
             "after re-open")


def test_small_block_repeats_bitwise(g):
    nj = 65536
    s = plummer_state(nj, seed=97)
    load_j(g, s)
    g.set_ti(0.0)
    xp, vp = _predict(s, 0.0)
    isel = np.array([3, 500, 40000, 65535])
    first = g.force(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8)
    for _ in range(3):
        again = g.force(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8)
        for u, w in zip(first, again):
            np.testing.assert_array_equal(u, w)


def _abi_block_run(g, s, nj, seed):
    """A kira-like sequence of ABI blocks: j updates carried in the launch
    arguments (at most 16) and by the scatter launch (17), blocks of 1 to 12
    i (one launch) and of 13 to 2048 i (the FP32 path, with j predicted in
    the force kernel up to 256 i, by the predictor above), blocks with
    neighbour spheres and their lists, and a j update made between the two
    halves of a block.  Returns every result in order."""
    rng = np.random.default_rng(seed)
    out = []
    ti = 0.0
    steps = [(1, 0), (2, 3), (4, 16), (8, 17), (3, 1), (5, 2), (12, 2), (64, 20), (300, 5),
             (2048, 0), (13, 5)]
    for step, (ni, nup) in enumerate(steps):
        ti += 2.0 ** -12
        for a in np.sort(rng.choice(nj, nup, replace=False)):
            s["x"][a] += rng.normal(scale=1e-3, size=3)
            s["tj"][a] = ti - 2.0 ** -13
            assert g.set_j_particle(int(a), int(s["idx"][a]), s["tj"][a], s["dtj"][a],
                                    s["m"][a], s["k18"][a], s["j6"][a], s["a2"][a], s["v"][a],
                                    s["x"][a]) == 0
        g.set_ti(ti)
        xp, vp = _predict(s, ti)
        isel = np.sort(rng.choice(nj, ni, replace=False))
        h2 = np.full(ni, 0.02) if step in (4, 7) else np.zeros(ni)
        g.calc_firsthalf(nj, s["idx"][isel], xp[isel], vp[isel], 1e-8, h2)
        if step in (5, 8):
            # the block in flight keeps the j-memory it started with
            a = int(rng.integers(nj))
            s["x"][a] += 1e-3
            assert g.set_j_particle(a, int(s["idx"][a]), s["tj"][a], s["dtj"][a], s["m"][a],
                                    s["k18"][a], s["j6"][a], s["a2"][a], s["v"][a],
                                    s["x"][a]) == 0
        rc, acc, jerk, pot, nn = g.calc_lasthalf2(nj, ni)
        assert rc == 0, g.last_error()
        out.append((acc, jerk, pot, nn))
        if step in (4, 7):
            assert g.read_neighbour_list() == 0
            for k in range(ni):
                rc, lst = g.get_neighbour_list(k, 4096)
                assert rc == 0
                out.append((lst,))
    return out


def test_completion_word_equals_event_wait(g):
    """lasthalf polls the completion word that a block's last launch (k_small,
    or the split reduction of the FP32 path) writes to pinned memory (option
    done_word = 1, the default) instead of waiting on a HIP event; every
    result of the same call sequence is the same bit for bit either way,
    lists and mid-block j updates included."""
    nj = 65536
    base = sphere_state(nj, seed=98)
    runs = {}
    try:
        for flag in (1, 0, 1):
            assert g.option("done_word", flag) == flag
            s = {k: v.copy() for k, v in base.items()}
            load_j(g, s)
            runs.setdefault(flag, []).append(_abi_block_run(g, s, nj, seed=6))
    finally:
        g.option("done_word", 1)
    for ref in runs[1][1:] + runs[0]:
        assert len(ref) == len(runs[1][0])
        for u, w in zip(runs[1][0], ref):
            for a, b in zip(u, w):
                np.testing.assert_array_equal(a, b)
