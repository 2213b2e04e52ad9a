"""GPU: the batched ensemble engine (include/g6ens.h) in shared-step mode
(option block_steps = 0; block mode is tested in test_gpu_ens_block.py).

* precision 64 with one j-split: every system's trajectory, step count and
  energies are bitwise those of the FP64 shared-step oracle
  (oracle.hermite_shared, same kira rules and operation order);
* precision 64 with j-splits: agreement to rounding (1e-12 of the system size
  over a short span);
* precision 32: the noise-floored step criterion keeps the step count within
  a small factor of the oracle's and energy conservation bounded;
* a system's result is bitwise identical alone and in a mix admitted at
  other times into other slots (continuous batching does not perturb it);
* admission by HBM budget: a pool too small for the whole mix still
  finishes every job, with jobs admitted as earlier ones retire.
"""
import numpy as np
import pytest

import oracle
from starlab_amd.ens import Ensemble, binary_single, plummer_cluster

pytestmark = pytest.mark.gpu

GB = 1 << 30


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


DEFAULTS = dict(jsplit_target=256, precision=64, max_steps=50_000_000, steps_per_graph=8,
                fused_steps=256, light_max=64, wg_steps=0, block_steps=1, heavy_lanes=1)
PATHS = pytest.mark.parametrize("fused", [256, 0], ids=["fused", "graphs"])


def _run(jobs, budget=2 * GB, max_systems=4096, **opt):
    opt.setdefault("block_steps", 0)
    e = Ensemble(budget, max_systems, **opt)
    try:
        ids = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in jobs]
        res = e.run_until_idle()
        stats = e.stats()
    finally:
        e.close()
        for k, v in DEFAULTS.items():           # engine options outlive the engine
            e.option(k, v)
    return [res[i] for i in ids], stats


@PATHS
def test_fp64_bitwise_equals_oracle(fused):
    """Light systems (n <= 64) run in the fused kernel or in step graphs."""
    jobs = _jobs(1, [3, 3, 16, 64, 200, 700, 1300], t_end=1.0 / 32)
    res, _ = _run(jobs, jsplit_target=0, fused_steps=fused)
    for j, r in zip(jobs, res):
        xf, vf, t, steps, e0, e1, rc = oracle.hermite_shared(j["m"], j["x"], j["v"], j["t_end"],
                                                            j["eta"], j["eps2"])
        print(len(j["m"]), "steps gpu/cpu", r["steps"], steps, "max|dx|",
              np.abs(r["x"] - xf).max(), flush=True)
        assert not r["failed"] and rc == 0
        assert r["t"] == t and r["steps"] == steps
        np.testing.assert_array_equal(r["x"], xf)
        np.testing.assert_array_equal(r["v"], vf)
        assert r["e0"] == e0 and r["e1"] == e1


def test_fp64_split_j_agrees_to_rounding():
    jobs = _jobs(8, [900, 2100], t_end=1.0 / 64)
    res, _ = _run(jobs, jsplit_target=256)
    for j, r in zip(jobs, res):
        xf, vf, t, steps, e0, e1, rc = oracle.hermite_shared(j["m"], j["x"], j["v"], j["t_end"],
                                                            j["eta"], j["eps2"])
        size = np.abs(j["x"]).max()
        assert r["steps"] == steps
        assert np.abs(r["x"] - xf).max() <= 1e-12 * size


def test_energy_conservation_vs_oracle():
    jobs = _jobs(2, [3, 32, 256, 1024], t_end=0.5)
    for prec in (64, 32):
        res, _ = _run(jobs, precision=prec)
        for j, r in zip(jobs, res):
            _, _, t, steps, e0, e1, rc = oracle.hermite_shared(j["m"], j["x"], j["v"], j["t_end"],
                                                              j["eta"], j["eps2"])
            de_gpu = abs((r["e1"] - r["e0"]) / r["e0"])
            de_cpu = abs((e1 - e0) / e0)
            print(prec, len(j["m"]), "dE/E gpu", de_gpu, "cpu", de_cpu, "steps", r["steps"],
                  steps, flush=True)
            assert not r["failed"] and r["t"] == 0.5
            if prec == 64:
                assert de_gpu <= 10 * de_cpu + 1e-12
            else:
                assert r["steps"] <= 4 * steps
                assert de_gpu <= 10 * de_cpu + 1e-5


@PATHS
@pytest.mark.parametrize("prec", [64, 32])
def test_alone_equals_mixed_bitwise(prec, fused):
    target = _jobs(3, [500], t_end=1.0 / 8)[0]
    target3 = _jobs(4, [3], t_end=2.0)[0]
    alone, _ = _run([target, target3], precision=prec, fused_steps=fused)
    mix = _jobs(5, [3, 64, 2048, 3, 700, 3, 128, 1500, 3, 3, 40], t_end=1.0 / 8)
    e = Ensemble(2 * GB, 64, precision=prec, fused_steps=fused, block_steps=0)
    try:
        others = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in mix[:6]]
        e.run(16)                                   # the mix is already running
        a = e.submit(target["m"], target["x"], target["v"], target["t_end"], target["eta"],
                     target["eps2"])
        more = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in mix[6:]]
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
    assert len(res) == len(mix) + 2


def test_admission_by_hbm_budget():
    jobs = _jobs(6, [1000] * 6 + [3] * 10, t_end=1.0 / 64)
    # room for ~2 clusters of 1000 at a time
    budget = 2500 * (21 * 8 + 64 + 7 * 8 * 16)
    res, stats = _run(jobs, budget=budget, max_systems=8)
    assert stats["finished"] == len(jobs)
    assert all(not r["failed"] and r["t"] == 1.0 / 64 for r in res)


def test_step_limit_marks_failure():
    job = _jobs(7, [64], t_end=1.0)[0]
    e = Ensemble(GB, 4, max_steps=5)
    try:
        jid = e.submit(job["m"], job["x"], job["v"], job["t_end"], job["eta"], job["eps2"])
        res = e.run_until_idle()
    finally:
        e.close()
        e.option("max_steps", 50_000_000)      # engine options outlive the engine
    assert res[jid]["failed"] and res[jid]["t"] < 1.0


@pytest.mark.parametrize("prec", [64, 32])
def test_fused_kernel_equals_step_graphs_bitwise(prec):
    """The fused light-system kernel performs the step kernels' arithmetic:
    same trajectories, step counts and energies bit for bit."""
    jobs = _jobs(9, [3, 3, 3, 5, 17, 40, 64], t_end=1.0)
    fused, _ = _run(jobs, precision=prec, fused_steps=256)
    graphs, _ = _run(jobs, precision=prec, fused_steps=0)
    for f, g in zip(fused, graphs):
        assert f["steps"] == g["steps"] and f["t"] == g["t"] == 1.0
        np.testing.assert_array_equal(f["x"], g["x"])
        np.testing.assert_array_equal(f["v"], g["v"])
        assert f["e0"] == g["e0"] and f["e1"] == g["e1"]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("jsplit", [0, 256])
def test_wg_kernel_equals_step_graphs_bitwise(prec, jsplit):
    """The workgroup-fused kernel (one workgroup per system of 64 < n <= 2048,
    whole steps per launch) performs the step kernels' arithmetic in their
    order, j-splits included: same trajectories, step counts and energies bit
    for bit, in both size classes (256 and 1024 threads per system) and beside
    a system too large for it (graph path) and light ones (wave-fused)."""
    jobs = _jobs(10, [65, 100, 300, 512, 513, 900, 2048, 2100, 3, 40], t_end=1.0 / 16)
    wg, _ = _run(jobs, precision=prec, jsplit_target=jsplit, wg_steps=32)
    graphs, _ = _run(jobs, precision=prec, jsplit_target=jsplit, wg_steps=0)
    for j, f, g in zip(jobs, wg, graphs):
        assert not f["failed"] and f["t"] == g["t"] == 1.0 / 16, len(j["m"])
        assert f["steps"] == g["steps"], len(j["m"])
        np.testing.assert_array_equal(f["x"], g["x"])
        np.testing.assert_array_equal(f["v"], g["v"])
        assert f["e0"] == g["e0"] and f["e1"] == g["e1"]


def test_wg_short_launches_resume_bitwise():
    """Stopping a workgroup-fused launch every step and resuming from the
    header gives the same trajectory as 32 steps per launch."""
    jobs = _jobs(11, [130, 700], t_end=1.0 / 32)
    a, _ = _run(jobs, precision=32, wg_steps=1)
    b, _ = _run(jobs, precision=32, wg_steps=32)
    for f, g in zip(a, b):
        assert f["steps"] == g["steps"]
        np.testing.assert_array_equal(f["x"], g["x"])
        assert f["e1"] == g["e1"]


def test_autotune_chooses_from_its_space():
    """The ensemble shape tuner times every candidate on copies of a sample
    and returns one of them; the jobs it was given are untouched."""
    from starlab_amd.ens import autotune
    jobs = _jobs(12, [3, 40, 300, 700, 1500], t_end=1.0 / 16)
    x0 = [j["x"].copy() for j in jobs]
    space = dict(jsplit_target=(256, 1024), wg_steps=(0, 32))
    try:
        best, timings, secs = autotune(jobs, sample=5, t_probe=1.0 / 256, budget=2 * GB,
                                       max_systems=64, space=space)
    finally:
        from starlab_amd import load
        L = load()
        for k, v in DEFAULTS.items():            # engine options outlive the engine
            L.g6e_set_option(k.encode(), int(v))
    assert best["jsplit_target"] in (256, 1024) and best["wg_steps"] in (0, 32)
    assert len(timings) == 4 and all(t > 0 for t in timings.values())
    assert secs > 0
    for j, x in zip(jobs, x0):
        np.testing.assert_array_equal(j["x"], x)



def test_failed_open_releases_and_bad_arguments_are_errors():
    """A g6e_open that fails after creating its streams (a budget beyond the
    pool's index range) frees them: the next open works.  NULL arrays to
    g6e_submit and g6e_poll are errors, and the engine carries on."""
    from starlab_amd import load
    L = load()
    assert L.g6e_open(1 << 62, 16) != 0
    assert b"too large" in L.g6e_last_error()
    eng = Ensemble(GB // 4, 16)
    try:
        assert eng.L.g6e_submit(3, None, None, None, 1.0, 0.1, 0.0) < 0
        assert b"null" in eng.L.g6e_last_error()
        assert eng.L.g6e_poll(4, None) < 0
        (job,) = _jobs(67, [3], 1.0 / 16)
        jid = eng.submit(job["m"], job["x"], job["v"], job["t_end"], job["eta"], job["eps2"])
        res = eng.run_until_idle()
        assert set(res) == {jid}
        assert res[jid]["t"] == job["t_end"] and not res[jid]["failed"]
    finally:
        eng.close()


def test_lane_options_fixed_while_open():
    """light_max, fused_steps, wg_steps and heavy_lanes decide which lane and kernel a
    resident system steps in: refused while an engine is open (a change
    under resident systems could leave a lane that never advances), taken
    before g6e_open, and a run with them set up front finishes."""
    from starlab_amd import load
    L = load()
    e = Ensemble(GB, 16, **DEFAULTS)
    try:
        for k in ("light_max", "fused_steps", "wg_steps", "heavy_lanes"):
            before = L.g6e_get_option(k.encode())
            assert L.g6e_set_option(k.encode(), 0) == -1
            assert b"fixed while open" in L.g6e_last_error()
            assert L.g6e_get_option(k.encode()) == before
        j = _jobs(5, [96], 1.0 / 64)[0]
        jid = e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"])
        res = e.run_until_idle()
        assert jid in res and res[jid]["t"] == j["t_end"] and not res[jid]["failed"]
    finally:
        e.close()
        for k, v in DEFAULTS.items():
            e.option(k, v)
    assert L.g6e_set_option(b"wg_steps", 0) == 0


def test_config4_sample_fp32_to_t1_against_fp64():
    """Config 4 at its real shape: a size-stratified sample of 256 of the 4096
    Plummer clusters (N ~ U[32, 2048], seed 4) integrated to t = 1 in FP32
    pairs, and the same systems in FP64 pairs (bitwise the oracle integrator,
    test_fp64_bitwise_equals_oracle).  Every system finishes.  The FP32 step
    rule subtracts a force-noise floor from kira's Aarseth criterion; against
    FP64 it takes within 6 % of the steps of every system (measured 0.984 to
    1.037, median 0.992), and its energy error stays below 2e-7 (measured
    max 8.6e-8, median 7.4e-9; FP64 max 4.3e-9, bound 1e-8).  Four of the systems are also
    run by the oracle itself on the CPU (systems of at most the j-split size,
    one j-split, so FP64 is bitwise the oracle): FP64 bitwise, FP32 within
    the same energy bound."""
    from starlab_amd import serve
    batch = serve.cluster_batch(4096, 32, 2048, 4, 1.0, prec=32)
    order = np.argsort([j["n"] for j in batch], kind="stable")
    sample = [batch[k] for k in order[::16]]
    r32 = serve.run_batch(sample, keep=True, block_steps=0)["results"]
    for j in sample:
        j["prec"] = 64
    r64 = serve.run_batch(sample, keep=True, block_steps=0)["results"]
    de32 = np.array([abs(r["e1"] / r["e0"] - 1) for r in r32])
    de64 = np.array([abs(r["e1"] / r["e0"] - 1) for r in r64])
    ratio = np.array([a["steps"] / b["steps"] for a, b in zip(r32, r64)])
    print(f"FP32 |dE/E| median {np.median(de32):.2e} max {de32.max():.2e}; FP64 median "
          f"{np.median(de64):.2e} max {de64.max():.2e}; steps FP32/FP64 median "
          f"{np.median(ratio):.3f} min {ratio.min():.3f} max {ratio.max():.3f}", flush=True)
    assert not any(r["failed"] for r in r32 + r64)
    assert all(r["t"] == 1.0 for r in r32 + r64)
    assert de32.max() <= 2e-7 and de64.max() <= 1e-8
    assert 0.94 <= ratio.min() and ratio.max() <= 1.06
    # one j-split (N at most the split size) sums every j in the oracle's order
    small = [k for k, j in enumerate(sample) if j["n"] <= DEFAULTS["jsplit_target"]]
    for k in small[::max(1, len(small) // 4)][:4]:
        j = sample[k]
        x, v, t, steps, e0, e1, rc = oracle.hermite_shared(j["m"], j["x"], j["v"], 1.0, j["eta"],
                                                           j["eps2"])
        assert rc == 0 and t == 1.0
        np.testing.assert_array_equal(r64[k]["x"], x)
        assert r64[k]["steps"] == steps
        assert abs(r32[k]["e1"] / r32[k]["e0"] - 1) <= 2e-7


@pytest.mark.parametrize("prec", [32, 64])
def test_alone_equals_mixed_at_the_largest_serving_cluster(prec):
    """The top of config 5's cluster range (N = 4096) gives the same result
    alone and admitted into a running mix of scatterings and clusters."""
    target = _jobs(13, [4096], t_end=1.0 / 64)[0]
    alone, _ = _run([target], precision=prec)
    mix = _jobs(14, [3, 3, 300, 3, 2048, 3, 3, 1000, 3], t_end=1.0 / 64)
    e = Ensemble(2 * GB, 64, precision=prec, block_steps=0)
    try:
        first = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in mix[:5]]
        e.run(8)
        a = e.submit(target["m"], target["x"], target["v"], target["t_end"], target["eta"],
                     target["eps2"])
        rest = [e.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"]) for j in mix[5:]]
        res = e.run_until_idle()
    finally:
        e.close()
        for k, v in DEFAULTS.items():
            e.option(k, v)
    np.testing.assert_array_equal(res[a]["x"], alone[0]["x"])
    np.testing.assert_array_equal(res[a]["v"], alone[0]["v"])
    assert res[a]["steps"] == alone[0]["steps"] and len(res) == len(mix) + 1


def test_poisson_serving_every_job_finishes_and_equals_alone():
    """A short served Poisson stream (2000 req/s, 400 jobs: 70 % three-body
    scatterings in FP64, 30 % clusters N ~ U[256, 4096] in FP32, admitted
    while others run): every job finishes, and a sample of them equals its
    run alone bit for bit."""
    from starlab_amd import serve
    jobs = serve.serving_mix(400, 21)
    r = serve.run_poisson(jobs, 2000.0, keep=True)
    print({k: v for k, v in r.items() if k != "results"}, flush=True)
    assert r["failed"] == 0 and len(r["results"]) == len(jobs)
    assert all(res["t"] == j["t_end"] for res, j in zip(r["results"], jobs))
    picks = [k for k, j in enumerate(jobs) if j["kind"] == "3body"][:4] + \
            [k for k, j in enumerate(jobs) if j["kind"] == "cluster"][:3]
    for k in picks:
        alone = serve.run_batch([jobs[k]], keep=True)["results"][0]
        np.testing.assert_array_equal(r["results"][k]["x"], alone["x"])
        assert r["results"][k]["steps"] == alone["steps"]
