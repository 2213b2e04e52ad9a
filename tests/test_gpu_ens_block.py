"""GPU: the ensemble engine in block mode (option block_steps = 1, the
default): kira's individual power-of-two block steps per system.

* precision 64 with one j-split: trajectory, block count and energies are
  bitwise those of the FP64 block-step oracle (oracle.hermite_block, pinned
  against kira's rules restated in starlab_amd/blockstep.py), on the step
  graphs and in the fused light-system kernel;
* with j-splits: agreement to rounding;
* the fused block path (prediction inside the force kernel, work lists built
  on the device, option block_fusion = 1, the default) is bitwise the four
  step kernels on predicted records (block_fusion = 0) in FP64, in both
  lanes, with and without j-splits.  In FP32 the fused path sums chunks of
  at most 32 i with several lanes per i (a different FP32 order), so there
  the two agree to FP32 rounding: positions within 1e-9 of the system size
  (measured up to 6e-11 against FP64 for either) and the same energy
  error to 1e-6.  Block counts are not compared between FP32 orders: kira's
  criterion differences successive forces, so over a short span FP32 noise
  can halve a few steps more in one order than the other (one 1300-body
  system took 665 blocks packed, 4837 unpacked, 694 in FP64).  Over config
  4 to t = 1 the FP32 block counts stay within 6 % of FP64 (last test);
* a system's result is bitwise the same alone and in a mix that holds
  block-step and shared-step systems of both precisions, admitted at other
  times;
* the pair interactions the engine counts are the oracle's particle steps;
* config 4's clusters at their real shape (a 256-system sample to t = 1):
  FP32 pairs against FP64 pairs, and block steps against shared steps.
"""
import numpy as np
import pytest

import oracle
from starlab_amd.ens import Ensemble, binary_single, plummer_cluster

pytestmark = pytest.mark.gpu

GB = 1 << 30
DEFAULTS = dict(jsplit_target=256, precision=64, max_steps=50_000_000, steps_per_graph=8,
                fused_steps=256, light_max=64, wg_steps=0, block_steps=1, block_fusion=1,
                wgb_steps=0, heavy_lanes=1)
PATHS = pytest.mark.parametrize("fused", [256, 0], ids=["fused", "graphs"])


def _jobs(seed, sizes, t_end, eta=0.05):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        if n == 3:
            m, x, v = binary_single(rng)
            eps2 = 0.0
        else:
            m, x, v = plummer_cluster(n, rng)
            eps2 = (4.0 / n) ** 2
        out.append(dict(m=m, x=x, v=v, t_end=t_end, eta=eta, eps2=eps2))
    return out


def _agree(f, g, prec, job, xtol=1e-9):
    """Two runs of one job: bitwise in FP64; in FP32 to the pair law's
    rounding (positions within xtol of the system size, relative energy
    errors within 1e-6 of each other: the FP32 potential differs in the last
    digits between sum orders, and a three-body scattering's total energy is
    a small difference of kinetic and potential energy)."""
    if prec == 64:
        assert f["steps"] == g["steps"], len(job["m"])
        np.testing.assert_array_equal(f["x"], g["x"])
        np.testing.assert_array_equal(f["v"], g["v"])
        assert f["e0"] == g["e0"] and f["e1"] == g["e1"]
    else:
        size = np.abs(job["x"]).max()
        print(len(job["m"]), "blocks", f["steps"], g["steps"], "max|dx|/size",
              np.abs(f["x"] - g["x"]).max() / size, flush=True)
        assert np.abs(f["x"] - g["x"]).max() <= xtol * size
        assert abs(f["e1"] / f["e0"] - g["e1"] / g["e0"]) <= 1e-6


def _run(jobs, budget=2 * GB, max_systems=4096, **opt):
    e = Ensemble(budget, max_systems, block_steps=1, **opt)
    try:
        ids = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in jobs]
        res = e.run_until_idle()
        stats = e.stats()
    finally:
        e.close()
        for k, v in DEFAULTS.items():           # engine options outlive the engine
            e.option(k, v)
    return [res[i] for i in ids], stats


def test_block_steps_are_the_default():
    from starlab_amd import load
    assert load().g6e_get_option(b"block_steps") == 1


@PATHS
def test_fp64_block_bitwise_equals_oracle(fused):
    jobs = _jobs(1, [3, 3, 16, 64, 200, 700, 1300], t_end=1.0 / 16)
    res, stats = _run(jobs, jsplit_target=0, fused_steps=fused)
    pairs = 0.0
    for j, r in zip(jobs, res):
        n = len(j["m"])
        xf, vf, t, nb, ps, e0, e1, rc = oracle.hermite_block(j["m"], j["x"], j["v"], j["t_end"],
                                                            j["eta"], j["eps2"])
        print(n, "blocks gpu/cpu", r["steps"], nb, "particle steps", ps, "max|dx|",
              np.abs(r["x"] - xf).max(), flush=True)
        assert not r["failed"] and rc == 0
        assert r["t"] == t == j["t_end"] and r["steps"] == nb
        np.testing.assert_array_equal(r["x"], xf)
        np.testing.assert_array_equal(r["v"], vf)
        assert r["e0"] == e0 and r["e1"] == e1
        pairs += (ps + 2 * n) * (n - 1)
        if n >= 200:
            assert ps < n * nb              # block steps, not one shared step
    assert stats["pairs"] == pairs


def test_fp64_block_split_j_agrees_to_rounding():
    jobs = _jobs(8, [900, 2100], t_end=1.0 / 64)
    res, _ = _run(jobs, jsplit_target=256)
    for j, r in zip(jobs, res):
        xf, vf, t, nb, ps, e0, e1, rc = oracle.hermite_block(j["m"], j["x"], j["v"], j["t_end"],
                                                            j["eta"], j["eps2"])
        size = np.abs(j["x"]).max()
        assert r["steps"] == nb
        assert np.abs(r["x"] - xf).max() <= 1e-12 * size


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("jsplit", [0, 256])
@pytest.mark.parametrize("fused", [256, 0], ids=["light-fused", "light-graphs"])
def test_fused_block_path_equals_step_kernels(prec, jsplit, fused):
    jobs = _jobs(15, [3, 40, 64, 65, 300, 700, 1300, 2100], t_end=1.0 / 16)
    a, sa = _run(jobs, precision=prec, jsplit_target=jsplit, fused_steps=fused, block_fusion=1)
    b, sb = _run(jobs, precision=prec, jsplit_target=jsplit, fused_steps=fused, block_fusion=0)
    for j, f, g in zip(jobs, a, b):
        assert not f["failed"] and f["t"] == g["t"] == 1.0 / 16, len(j["m"])
        _agree(f, g, prec, j)
    if prec == 64:
        assert sa["pairs"] == sb["pairs"]


@pytest.mark.parametrize("prec", [64, 32])
def test_fused_light_kernel_equals_step_graphs(prec):
    """Light systems in the fused light-system kernel against the same
    systems on the block path's step graphs: bitwise in FP64, to FP32
    rounding in FP32 (where the graphs pack several lanes per i; over a whole
    time unit of three-body encounters positions agree to 1e-7 of the system
    size, measured 4e-9)."""
    jobs = _jobs(9, [3, 3, 3, 5, 17, 40, 64], t_end=1.0)
    fused, _ = _run(jobs, precision=prec, fused_steps=256)
    graphs, _ = _run(jobs, precision=prec, fused_steps=0)
    for j, f, g in zip(jobs, fused, graphs):
        assert f["t"] == g["t"] == 1.0
        _agree(f, g, prec, j, xtol=1e-7)


@PATHS
@pytest.mark.parametrize("prec", [64, 32])
def test_alone_equals_mixed_bitwise(prec, fused):
    """The mix holds systems of the other step mode and the other precision
    as well, admitted before and after the targets."""
    target = _jobs(3, [500], t_end=1.0 / 8)[0]
    target3 = _jobs(4, [3], t_end=2.0)[0]
    alone, _ = _run([target, target3], precision=prec, fused_steps=fused)
    mix = _jobs(5, [3, 64, 2048, 3, 700, 3, 128, 1500, 3, 3, 40], t_end=1.0 / 8)
    e = Ensemble(2 * GB, 64, precision=prec, fused_steps=fused, block_steps=1)
    try:
        others = []
        for k, j in enumerate(mix[:6]):
            e.option("block_steps", k % 2)
            e.option("precision", 64 if k % 3 else 32)
            others.append(e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]))
        e.run(16)                                   # the mix is already running
        e.option("block_steps", 1)
        e.option("precision", prec)
        a = e.submit(target["m"], target["x"], target["v"], target["t_end"], target["eta"],
                     target["eps2"])
        for k, j in enumerate(mix[6:]):
            e.option("block_steps", (k + 1) % 2)
            e.option("precision", 32 if k % 2 else 64)
            others.append(e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]))
        e.option("block_steps", 1)
        e.option("precision", prec)
        b = e.submit(target3["m"], target3["x"], target3["v"], target3["t_end"], target3["eta"],
                     target3["eps2"])
        res = e.run_until_idle()
    finally:
        e.close()
        for k, v in DEFAULTS.items():
            e.option(k, v)
    for ref, got in ((alone[0], res[a]), (alone[1], res[b])):
        np.testing.assert_array_equal(got["x"], ref["x"])
        np.testing.assert_array_equal(got["v"], ref["v"])
        assert got["steps"] == ref["steps"] and got["e1"] == ref["e1"]
    assert len(res) == len(mix) + 2 and not any(r["failed"] for r in res.values())


@pytest.mark.parametrize("prec", [64, 32])
def test_two_heavy_lanes_equal_one_bitwise(prec):
    """Option heavy_lanes = 2 deals the heavy systems (n > light_max) to two
    lanes, each with its own stream and graphs.  A system's arithmetic does
    not depend on the lane that steps it: results are bitwise those of one
    heavy lane, for block-step and shared-step systems, with j-splits, beside
    light systems, and the engine counts the same pairs."""
    jobs = _jobs(17, [3, 40, 65, 100, 300, 700, 1300, 2100, 150, 520], t_end=1.0 / 16)
    out = []
    for lanes in (1, 2):
        e = Ensemble(2 * GB, 64, precision=prec, heavy_lanes=lanes, block_steps=1)
        try:
            ids = []
            for k, j in enumerate(jobs):
                e.option("block_steps", 0 if k % 3 == 2 else 1)
                ids.append(e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]))
            res = e.run_until_idle()
            stats = e.stats()
            assert e.option("heavy_lanes") == lanes
        finally:
            e.close()
            for k, v in DEFAULTS.items():
                e.option(k, v)
        out.append(([res[i] for i in ids], stats))
    (one, s1), (two, s2) = out
    for j, f, g in zip(jobs, one, two):
        assert not f["failed"] and f["t"] == g["t"] == 1.0 / 16, len(j["m"])
        assert f["steps"] == g["steps"], len(j["m"])
        np.testing.assert_array_equal(f["x"], g["x"])
        np.testing.assert_array_equal(f["v"], g["v"])
        assert f["e0"] == g["e0"] and f["e1"] == g["e1"]
    assert s1["pairs"] == s2["pairs"]


def test_step_limit_marks_failure():
    job = _jobs(7, [64], t_end=1.0)[0]
    res, _ = _run([job], budget=GB, max_systems=4, max_steps=5)
    assert res[0]["failed"] and res[0]["t"] < 1.0


def test_config4_sample_to_t1_block_fp32_fp64_and_shared():
    """Config 4 at its real shape: every 16th of the 4096 Plummer clusters
    (N ~ U[32, 2048], seed 4) by size, to t = 1, on block steps in FP32 and
    in FP64 pairs, and on shared steps in FP32.  Every system finishes.
    Against FP64 the FP32 block runs take within 6 % of the blocks (measured
    0.984 to 1.005), and their energy error stays below 2e-7 (measured max
    8.3e-8, median 8.6e-9).  The FP64 block runs' own error, the scheme's at
    eta = 0.1 with steps longer than the shared ones, stays below 1e-7
    (measured max 3.8e-8, median 5.5e-9).  Block steps evaluate at least 10x
    fewer pair interactions than shared steps over the sample (measured
    22.7x).
    Four small systems (N at most the j-split size) are also run by the
    block-step oracle on the CPU (one j-split, so FP64 is bitwise the
    oracle)."""
    from starlab_amd import serve
    batch = serve.cluster_batch(4096, 32, 2048, 4, 1.0, prec=32)
    order = np.argsort([j["n"] for j in batch], kind="stable")
    sample = [batch[k] for k in order[::16]]
    b32 = serve.run_batch(sample, keep=True, block_steps=1)
    s32 = serve.run_batch(sample, keep=False, block_steps=0)
    for j in sample:
        j["prec"] = 64
    b64 = serve.run_batch(sample, keep=True, block_steps=1)
    r32, r64 = b32["results"], b64["results"]
    de32 = np.array([abs(r["e1"] / r["e0"] - 1) for r in r32])
    de64 = np.array([abs(r["e1"] / r["e0"] - 1) for r in r64])
    ratio = np.array([a["steps"] / b["steps"] for a, b in zip(r32, r64)])
    print(f"block FP32 |dE/E| median {np.median(de32):.2e} max {de32.max():.2e}; FP64 median "
          f"{np.median(de64):.2e} max {de64.max():.2e}; blocks FP32/FP64 median "
          f"{np.median(ratio):.3f} min {ratio.min():.3f} max {ratio.max():.3f}; pairs block "
          f"{b32['pairs']:.3e} shared {s32['pairs']:.3e}; wall block {b32['wall_s']:.2f} s "
          f"shared {s32['wall_s']:.2f} s", flush=True)
    assert b32["failed"] == b64["failed"] == s32["failed"] == 0
    assert all(r["t"] == 1.0 for r in r32 + r64)
    assert de32.max() <= 2e-7 and de64.max() <= 1e-7
    assert 0.94 <= ratio.min() and ratio.max() <= 1.06
    assert s32["pairs"] >= 10 * b32["pairs"]
    # one j-split (N at most the split size) sums every j in the oracle's order
    small = [k for k, j in enumerate(sample) if j["n"] <= DEFAULTS["jsplit_target"]]
    for k in small[::max(1, len(small) // 4)][:4]:
        j = sample[k]
        x, v, t, nb, ps, e0, e1, rc = oracle.hermite_block(j["m"], j["x"], j["v"], 1.0, j["eta"],
                                                          j["eps2"])
        assert rc == 0 and t == 1.0
        np.testing.assert_array_equal(r64[k]["x"], x)
        assert r64[k]["steps"] == nb


def test_kernel_timing_runs_the_same_steps():
    """Option kernel_timing launches the step kernels one at a time between
    HIP events instead of as graphs: the same results bit for bit, and a
    device time for every kernel of the block path and the light lane."""
    jobs = _jobs(16, [3, 40, 300, 1300], t_end=1.0 / 16)
    a, _ = _run(jobs, precision=32)
    e = Ensemble(2 * GB, 64, block_steps=1, precision=32, kernel_timing=1)
    try:
        ids = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in jobs]
        res = e.run_until_idle()
        us = {k: e.option("kernel_us_" + k) for k in ("bselect", "bitems", "bforce", "bcorrect",
                                                       "finalize", "fused")}
    finally:
        e.close()
        e.option("kernel_timing", 0)
        for k, v in DEFAULTS.items():
            e.option(k, v)
    print(us, flush=True)
    for r0, i in zip(a, ids):
        np.testing.assert_array_equal(res[i]["x"], r0["x"])
        assert res[i]["steps"] == r0["steps"]
    assert all(v > 0 for v in us.values())


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("jsplit", [0, 256])
def test_workgroup_block_kernel_equals_block_path_bitwise(prec, jsplit):
    """k_ens_wgb (one workgroup per system of 64 < n <= 2048, block steps
    inside the launch) performs the fused block path's operations on the same
    values: trajectories, block counts and energies bit for bit, in both
    thread counts (256 up to 512 bodies, 512 above), beside a system too
    large for it (block path) and light ones (wave-fused kernel)."""
    jobs = _jobs(17, [65, 100, 300, 512, 513, 900, 2048, 2100, 3, 40], t_end=1.0 / 16)
    wg, sw = _run(jobs, precision=prec, jsplit_target=jsplit, wgb_steps=32)
    bp, sb = _run(jobs, precision=prec, jsplit_target=jsplit, wgb_steps=0)
    for j, f, g in zip(jobs, wg, bp):
        assert not f["failed"] and f["t"] == g["t"] == 1.0 / 16, len(j["m"])
        assert f["steps"] == g["steps"], len(j["m"])
        np.testing.assert_array_equal(f["x"], g["x"])
        np.testing.assert_array_equal(f["v"], g["v"])
        assert f["e0"] == g["e0"] and f["e1"] == g["e1"]
    assert sw["pairs"] == sb["pairs"]


def test_workgroup_block_kernel_short_launches_resume_bitwise():
    """One block step per launch, resumed from the header and the pool,
    gives the trajectory of 64 steps per launch."""
    jobs = _jobs(18, [130, 700, 1500], t_end=1.0 / 32)
    a, _ = _run(jobs, precision=32, wgb_steps=1)
    b, _ = _run(jobs, precision=32, wgb_steps=64)
    for f, g in zip(a, b):
        assert f["steps"] == g["steps"]
        np.testing.assert_array_equal(f["x"], g["x"])
        assert f["e1"] == g["e1"]


@PATHS
def test_block_edge_cases_bitwise_equal_oracle(fused):
    """Block steps at the edges, FP64 against the block-step oracle bit for
    bit: an end time that is no multiple of the steps (each particle's last
    step is cut to end there), one body alone, two bodies, and a light
    system beside a medium one."""
    rng = np.random.default_rng(19)
    jobs = []
    m, x, v = plummer_cluster(90, rng)
    jobs.append(dict(m=m, x=x, v=v, t_end=0.3, eta=0.05, eps2=(4.0 / 90) ** 2))
    jobs.append(dict(m=np.array([1.0]), x=np.array([[0.1, 0.2, 0.3]]), v=np.array([[0.0, 0.1, 0.0]]),
                     t_end=0.25, eta=0.05, eps2=0.0))
    jobs.append(dict(m=np.array([0.7, 0.3]), x=np.array([[-0.3, 0, 0], [0.7, 0, 0]]),
                     v=np.array([[0, -0.25, 0], [0, 0.6, 0]]), t_end=1.7, eta=0.02, eps2=0.0))
    m, x, v = plummer_cluster(20, rng)
    jobs.append(dict(m=m, x=x, v=v, t_end=0.3, eta=0.05, eps2=(4.0 / 20) ** 2))
    res, _ = _run(jobs, jsplit_target=0, fused_steps=fused)
    for j, r in zip(jobs, res):
        xf, vf, t, nb, ps, e0, e1, rc = oracle.hermite_block(j["m"], j["x"], j["v"], j["t_end"],
                                                            j["eta"], j["eps2"])
        assert rc == 0 and not r["failed"], len(j["m"])
        assert r["t"] == t == j["t_end"] and r["steps"] == nb, len(j["m"])
        np.testing.assert_array_equal(r["x"], xf)
        np.testing.assert_array_equal(r["v"], vf)
        assert r["e1"] == e1
