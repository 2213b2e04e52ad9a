#!/usr/bin/env python3
"""Benchmark: the GRAPE-6 force path of kira on one MI355X.

Workload (force-only benchmark): Plummer N = 65536, equal masses, N-body
units, eps^2 = 1e-8, every i against every j, j predicted from times t_j
spread uniformly over [t_i - 2^-10, t_i].  One step = one full sweep on the
device: j predictor + force/jerk/pot/nearest-neighbour kernel + split
reduction + FP64 nearest-neighbour refinement, inputs resident in HBM.

Prints one JSON line: pair interactions per second (N*(N-1)*steps / time),
the roofline of the force kernel against the FP32 vector peak (60 flop per
pair interaction), the ensemble legs (batched clusters; a served Poisson
stream with p50/p99 latency), kira end to end (plain and with kira's whole
GRAPE-6 call mix), the CPU direct-sum baselines measured on this host, and
the error against FP64 for every configuration (error_vs_fp64).  The GPU
legs keep samples of their inputs and results; the FP64 references are
computed by the CPU leg, outside every timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 65536
EPS2 = 1e-8
TSPREAD = 2.0 ** -10
FLOP_PER_PAIR = 60
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak (AMD published)
SEED = 65536


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_input(g, n):
    """Plummer sphere (scaled with a device pot sweep) plus predictor inputs
    (a, jerk from a device sweep at t = 0; snap = 0)."""
    from starlab_amd.ics import makeplummer
    idx = np.arange(n, dtype=np.int32)
    z = np.zeros((n, 3))

    def dev_potential(m, x):
        g.set_j_particles(idx, idx, np.zeros(n), np.ones(n), m, z, z, z, z, x)
        assert g.self_sweep(n, 0, n, 0.0, 0.0) == 0
        _, _, pot, _ = g.read_results(n)
        return 0.5 * float((m * pot).sum())

    m, x, v = makeplummer(n, SEED, potential_fn=dev_potential)
    g.set_j_particles(idx, idx, np.zeros(n), np.ones(n), m, z, z, z, v, x)
    assert g.self_sweep(n, 0, n, 0.0, EPS2) == 0
    acc, jerk, _, _ = g.read_results(n)
    rng = np.random.default_rng(SEED)
    tj = -rng.uniform(0.0, TSPREAD, n)
    g.set_j_particles(idx, idx, tj, np.full(n, TSPREAD), m, z, jerk / 6.0, acc / 2.0, v, x)
    return dict(m=m, x=x, v=v, a2=acc / 2.0, j6=jerk / 6.0, tj=tj, idx=idx)


def _pct(a):
    a = np.asarray(a, dtype=np.float64)
    if a.size == 0:
        return None
    return {"median": float(np.median(a)), "p99": float(np.percentile(a, 99)),
            "max": float(a.max())}


def fp64_errors(samples, eps2):
    """Error of kept GPU results against the FP64 oracle on the same inputs:
    per-particle |da|/|a|, |dj|/|j|, |dpot/pot| percentiles, the normwise
    acc / jerk error (|da| over sum_j |a_ij|), nearest-neighbour mismatches.
    Each sample: the j-memory as the library held it (t, m, x, v, acc, jerk,
    k18; predicted here to t), the i inputs and the library's results."""
    import oracle
    ra_, rj_, rp_, na_, nj_ = [], [], [], [], []
    mism = 0
    for s in samples:
        jm = s["jm"]
        n = len(jm["m"])
        jidx = np.arange(n, dtype=np.int32)
        a2 = jm["a2"] if "a2" in jm else jm["acc"] / 2.0     # as sent to the library
        j6 = jm["j6"] if "j6" in jm else jm["jerk"] / 6.0
        xp, vp = oracle.predict(jm["x"], jm["v"], a2, j6, jm["k18"], jm["t"], s["t"], 4)
        racc, rjerk, rpot, rnn, _ = oracle.force(s["xi"], s["vi"], s["idx"], xp, vp, jm["m"],
                                                 jidx, eps2)
        for k, i in enumerate(s["idx"]):
            d = xp - s["xi"][k]
            w = vp - s["vi"][k]
            r2 = (d * d).sum(1) + eps2
            if 0 <= i < n:
                r2[i] = np.inf
            sa = (jm["m"] / r2).sum()
            sj = (jm["m"] * 4.0 * np.sqrt((w * w).sum(1)) / (r2 * np.sqrt(r2))).sum()
            da = np.linalg.norm(s["acc"][k] - racc[k])
            dj = np.linalg.norm(s["jerk"][k] - rjerk[k])
            ra_.append(da / np.linalg.norm(racc[k]))
            rj_.append(dj / np.linalg.norm(rjerk[k]))
            rp_.append(abs(s["pot"][k] - rpot[k]) / abs(rpot[k]))
            na_.append(da / sa)
            nj_.append(dj / sj)
        mism += int((np.asarray(s["nn"]) != rnn).sum())
    return {"particle_steps": len(ra_), "acc_rel": _pct(ra_), "jerk_rel": _pct(rj_),
            "pot_rel": _pct(rp_), "acc_normwise": _pct(na_), "jerk_normwise": _pct(nj_),
            "nn_mismatch": mism}


def cpu_leg(inp, gpu, target_s=10.0, clusters=None, kira=None):
    """CPU side, outside every timed GPU region.

    Baseline (config 2's workload): the FP64 direct sum + nearest neighbour of
    oracle/g6_oracle.c built for speed (-O3, FMA, OpenMP over i, x86-64-v4 on
    AVX-512 hosts else v3) on all host threads, on a bounded sample of i
    against all N j, ~target_s of work; the bit-exact oracle build (-O2, no
    contraction) is timed beside it.  Config 3: the kira leg's pair
    interactions at that rate (the CPU force path of kira, one N-body unit).
    Config 4: a size-stratified sample of the ensemble's clusters integrated
    with the same block-step Hermite rules in FP64 (oracle.hermite_block),
    one system per host thread.  Errors against FP64 for configs 1-3 from the
    kept GPU samples."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    n = len(inp["m"])
    xp, vp = oracle.predict(inp["x"], inp["v"], inp["a2"], inp["j6"], np.zeros((n, 3)),
                            inp["tj"], 0.0, 4)
    idx = inp["idx"]
    L, flags = oracle.fast_lib()
    model, _ = oracle.host_cpu()

    def force_with(lib, ni):
        acc = np.zeros((ni, 3)); jerk = np.zeros((ni, 3)); pot = np.zeros(ni)
        nn = np.zeros(ni, np.int32); d2 = np.zeros(ni)
        lib.or_force(ni, oracle._d(xp[:ni]), oracle._d(vp[:ni]), oracle._i(idx[:ni]), n,
                     oracle._d(xp), oracle._d(vp), oracle._d(inp["m"]), oracle._i(idx), EPS2,
                     acc, jerk, pot, nn, d2)

    def rate(lib, budget):
        t0 = time.perf_counter()
        force_with(lib, 256)
        dt = time.perf_counter() - t0
        ni = int(min(n, max(256, 256 * budget / max(dt, 1e-6))))
        reps = 1 if ni < n else max(1, min(8, int(budget / max(dt * n / 256, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(reps):
            force_with(lib, ni)
        dt = time.perf_counter() - t0
        return reps * ni * (n - 1) / dt, reps, ni, dt

    value, reps, ni, dt = rate(L, target_s)
    strict, _, _, _ = rate(oracle.lib(), target_s / 4)
    out = {"value": value, "unit": "pair interactions/s", "cores": oracle.threads(),
           "kind": "port", "seconds": round(dt, 3), "cpu_model": model, "build": flags,
           "sample": f"{reps} x {ni} i-particles x {n} j of config 2 (FP64 direct sum + "
                     f"nearest neighbour, oracle/g6_oracle.c, OpenMP over i)",
           "oracle_build_value": strict}
    if kira is not None:
        pairs = kira["particle_steps"] * (kira["n"] - 1)
        out["config3"] = {"pairs": pairs, "force_seconds_at_cpu_rate": pairs / value,
                          "per_nbody_unit_s": pairs / value / kira["t_end"],
                          "gpu_end_to_end_s": kira["seconds"],
                          "note": "the kira leg's pair interactions at the CPU rate above "
                                  "(kira's host force path for the same blocks)"}
    if clusters:
        threads = oracle.threads()
        sample = clusters["sample"]

        def one(j):
            # pair interactions counted as the engine counts them: each
            # particle step, plus every particle at the start and the end
            n = len(j["m"])
            _, _, _, _, psteps, _, _, _ = oracle.hermite_block(j["m"], j["x"], j["v"], j["t_end"],
                                                               j["eta"], j["eps2"], L=L)
            return (psteps + 2 * n) * (n - 1)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            pairs = sum(ex.map(one, sample))
        dt4 = time.perf_counter() - t0
        out["config4"] = {"pairs_per_s": pairs / dt4, "systems_per_s": len(sample) / dt4,
                          "seconds": round(dt4, 3), "threads": threads,
                          "sample": f"{len(sample)} of the 4096 clusters (every "
                                    f"{clusters['stride']}th), FP64, block steps, to "
                                    f"t = {sample[0]['t_end']}",
                          "gpu_batch_pairs_per_s": clusters.get("gpu_pairs_per_s"),
                          "gpu_batch_t1_s_at_cpu_rate":
                              clusters["gpu_pairs_t1"] / (pairs / dt4)
                              if clusters.get("gpu_pairs_t1") else None}
    errs = {}
    sel = np.arange(0, n, max(1, n // 512))[:512]
    errs["config2"] = fp64_errors([{"t": 0.0, "jm": gpu["jm"], "idx": idx[sel], "xi": xp[sel],
                                    "vi": vp[sel], "acc": gpu["acc"][sel],
                                    "jerk": gpu["jerk"][sel], "pot": gpu["pot"][sel],
                                    "nn": gpu["nn"][sel]}], EPS2)
    for name, samples, eps2 in gpu.get("samples", ()):
        errs[name] = fp64_errors(samples, eps2)
    return out, errs


def ensemble_legs(seconds=2.0):
    """Batched ensembles and the served stream.

    batch_fp32_t1: 4096 Plummer clusters, N ~ U[32, 2048], FP32 pairs,
    integrated to t = 1 on kira's block steps (the batched-ensemble
    configuration as specified); batch_fp32_t1_shared: the same on one shared
    step per system (every particle moves at every step); batch_fp64: the
    clusters in FP64 to t = 1/32 on block steps; all with engine shapes from
    the ensemble autotuner.  serving: a Poisson stream of 2000 req/s, 70 %
    binary-single scatterings (FP64) to t = 32, 30 % clusters N ~ U[256,
    4096] (FP32) to t = 1/32, on block steps."""
    from starlab_amd import serve
    keys = ("jobs", "wall_s", "systems_per_s", "pairs", "pairs_per_s", "engine_steps", "p50_s",
            "p99_s", "de_median", "de_max", "failed", "tuned")
    out = {}
    for name, prec, t_end, block in (("batch_fp32_t1", 32, 1.0, 1),
                                     ("batch_fp32_t1_shared", 32, 1.0, 0),
                                     ("batch_fp64", 64, 1.0 / 32, 1)):
        batch = serve.cluster_batch(4096, 32, 2048, 4, t_end, prec=prec)
        r = serve.run_batch(batch, tune=True, block_steps=block)
        out[name] = {k: r[k] for k in keys}
        out[name]["t_end"] = t_end
        out[name]["steps"] = "block" if block else "shared"
    jobs = serve.serving_mix(int(2000 * seconds), 5)
    r = serve.run_poisson(jobs, 2000.0)
    out["serving"] = r
    # error against FP64: the same 256 clusters (every 16th of the batch) in
    # both precisions to t = 1/32 on block steps; the FP64 engine is the
    # block-step oracle integrator bit for bit (tests/test_gpu_ens_block.py)
    sample = serve.cluster_batch(4096, 32, 2048, 4, 1.0 / 32, prec=64)[::16]
    r64 = serve.run_batch(sample, keep=True)["results"]
    for j in sample:
        j["prec"] = 32
    r32 = serve.run_batch(sample, keep=True)["results"]
    dx, de32, de64 = [], [], []
    for a, b in zip(r32, r64):
        scale = np.sqrt((b["x"] ** 2).sum(1).mean())
        dx.append(float(np.abs(a["x"] - b["x"]).max() / scale))
        de32.append(abs(a["e1"] / a["e0"] - 1))
        de64.append(abs(b["e1"] / b["e0"] - 1))
    out["error_vs_fp64"] = {
        "systems": len(sample), "t_end": 1.0 / 32,
        "position_rel_max_fp32_vs_fp64": {"median": float(np.median(dx)),
                                          "p99": float(np.percentile(dx, 99)),
                                          "max": float(np.max(dx))},
        "de_fp32": {"median": float(np.median(de32)), "max": float(np.max(de32))},
        "de_fp64": {"median": float(np.median(de64)), "max": float(np.max(de64))},
        "force": "FP32 pairs on double-single positions: the pair law of config 2"}
    return out


def aux_legs(g, inp, reps=3):
    """The energy and density fast paths on the same j-memory (include/g6x.h:
    g6x_potential, pot only; g6x_densities, k = 12 nearest neighbours), each
    timed over `reps` calls after one warm call, with its kernel time."""
    n = len(inp["m"])
    out = {}
    g.option("timing", 1)
    for name, kern, fn in (("potential", "pot", lambda: g.potential(n, 0, n, 0.0, EPS2)),
                           ("densities_k12", "knn", lambda: g.densities(n, 0, n, 0.0, 12))):
        fn()
        g.reset_timing()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        wall = (time.perf_counter() - t0) / reps
        calls, ms = g.kernel_time(kern)
        out[name] = {"wall_ms": round(wall * 1e3, 3),
                     "kernel_ms": round(ms / max(calls, 1), 4),
                     "pairs_per_s_kernel": n * (n - 1) / (ms / max(calls, 1) * 1e-3)}
    g.option("timing", 0)
    return out


def grape4_leg(inp, xp, vp, npipe=2048, sweeps=7):
    """Full sweeps through the GRAPE-4 ABI (libg4x.so) the way kira drives it
    with USE_HALF_CHIP: every particle in two adjacent pipes, 1024 particles per
    h3calc call, host buffers in and out (PCIe and host packing included).
    xp, vp: the j-particles predicted to t = 0 (read back from libg6), which
    kira would have predicted on the host.  A sweep takes ~10 ms, so one host
    stall can dominate it: the median of `sweeps` sweeps is reported."""
    from starlab_amd.g4 import GRAPE4
    n = len(inp["m"])
    g4 = GRAPE4()
    g4.open()
    g4.load_all(inp["x"], inp["v"], inp["a2"], inp["j6"], inp["m"], inp["tj"])
    g4.setti(0.0)
    per = npipe // 2
    blocks = [np.repeat(np.arange(s, min(n, s + per)), 2) for s in range(0, n, per)]
    g4.calc(n, xp[blocks[0]], vp[blocks[0]], EPS2 / 2, 0.0)      # warm
    walls = []
    for _ in range(sweeps):
        t0 = time.perf_counter()
        for b in blocks:
            g4.calc(n, xp[b], vp[b], EPS2 / 2, 0.0)
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    slots = g4.option("last_slots")
    g4.close()
    return {"pairs_per_s": n * (n - 1) / wall, "wall_ms": round(wall * 1e3, 3),
            "wall_ms_min": round(min(walls) * 1e3, 3), "wall_ms_max": round(max(walls) * 1e3, 3),
            "sweeps": sweeps, "calls": len(blocks), "pipes_per_call": npipe,
            "distinct_i_per_call": slots}


def abi_leg(nj=N, blocks=200, sizes=(1, 16, 256, 2048), h2=0.0):
    """kira's per-block sequence through the GRAPE-6 ABI (set_j of the previous
    block, set_ti, firsthalf, lasthalf2) from the compiled C caller
    starlab_amd/abi_bench: microseconds per block, host-to-host.  With h2 > 0
    every i asks for its neighbour list (radius^2 h2) and each block also
    runs g6_read_neighbour_list_ and g6_get_neighbour_list_ per i."""
    exe = os.path.join(ROOT, "starlab_amd", "abi_bench")
    env = dict(os.environ)
    if h2 > 0:
        env["ABI_BENCH_H2"] = repr(h2)
    r = subprocess.run([exe, str(nj), str(blocks), *map(str, sizes)], capture_output=True,
                       text=True, timeout=300, env=env)
    if r.returncode:
        raise RuntimeError(f"abi_bench failed ({r.returncode}): {r.stderr.strip()}")
    keys = ["ni", "us_per_block", "us_firsthalf", "us_lasthalf", "pairs_per_s"]
    if h2 > 0:
        keys += ["us_lists", "mean_list_length"]
    return {"nj": nj, "blocks": blocks, "h2": h2,
            "per_block": [{k: json.loads(line)[k] for k in keys} for line in r.stdout.splitlines()]}


def _keeper(samples, every, max_i, seed):
    """Observer for NativeBlockHermite keeping every `every`-th force call's
    j-memory, i inputs and results (at most max_i i, seeded) for the error
    against FP64; returns (observer, seconds spent in it)."""
    rng = np.random.default_rng(seed)
    calls = [0]
    spent = [0.0]

    def obs(jm, t, idx, xi, vi, acc, jerk, pot, nn):
        calls[0] += 1
        if (calls[0] - 1) % every:
            return False
        t0 = time.perf_counter()
        sel = np.arange(len(idx))
        if max_i and len(idx) > max_i:
            sel = np.sort(rng.choice(len(idx), max_i, replace=False))
        samples.append(dict(t=t, jm={k: np.array(jm[k]) for k in jm}, idx=np.array(idx[sel]),
                            xi=np.array(xi[sel]), vi=np.array(vi[sel]), acc=np.array(acc[sel]),
                            jerk=np.array(jerk[sel]), pot=np.array(pot[sel]),
                            nn=np.array(nn[sel])))
        spent[0] += time.perf_counter() - t0
        return False

    return obs, spent


def config1_leg(samples, every=23):
    """Config 1: Plummer N = 1024 (makeplummer -s 42), eps = 0, eta = 0.1,
    t = 0 -> 1 through the native block-step driver on libg6; every
    `every`-th block is kept for the error against FP64."""
    from starlab_amd.blockstep import NativeBlockHermite
    from starlab_amd.ics import makeplummer
    m, x, v = makeplummer(1024, 42)
    obs, spent = _keeper(samples, every, None, 1)
    b = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(1.0, observer=obs)
    return {"blocks": b.blocks, "particle_steps": b.particle_steps,
            "seconds": round(b.seconds - spent[0], 3), "kept_blocks": len(samples)}


def kira_leg(n=262144, t_end=0.25, samples=None):
    """kira end to end on libg6 (config 3's workload): King W0 = 6, N = 262144
    top-level nodes, Salpeter 1-100, 5 % binaries as centres of mass, eta =
    0.1, eps = 0, t = 0 -> 1/4, integrated by the native block-step driver
    (libkira_blockstep.so) with one g6_set_j_particle_ per particle as kira
    makes them.  Energies from g6x_potential (FP32 pairs, FP64 sums)."""
    from starlab_amd import ics
    from starlab_amd.blockstep import NativeBlockHermite
    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    if g.open() != 0:
        raise RuntimeError(g.last_error())
    idx = np.arange(n, dtype=np.int32)
    z = np.zeros((n, 3))

    def pot_energy(m, x):
        g.set_j_particles(idx, idx, np.zeros(n), np.ones(n), m, z, z, z, z, x)
        return 0.5 * float((m * g.potential(n, 0, n, 0.0, 0.0)).sum())

    def energy(m, x, v):
        return pot_energy(m, x) + 0.5 * float((m * (v * v).sum(1)).sum())

    c = ics.king_cluster(n, 42, potential_fn=pot_energy)
    m, x, v = c["m"], c["x"], c["v"]
    e0 = energy(m, x, v)
    g.close()
    b = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(t_end)
    g = GRAPE6()
    if g.open() != 0:
        raise RuntimeError(g.last_error())
    e1 = energy(b.m, b.x, b.v)
    g.close()
    wall = b.seconds
    if samples is not None:
        # a separate run keeps samples, so the timed one calls no observer:
        # every 197th block of the first 1/32 time unit, at most 16 i each
        obs, _ = _keeper(samples, 197, 16, 3)
        NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(t_end / 8, observer=obs)
    return {"workload": f"King W0=6 N={n}, Salpeter 1-100, 5% binaries (centres of mass), "
                        f"eta=0.1, eps=0, t=0->{t_end}; native kira block-step driver, "
                        "per-particle g6_set_j_particle_",
            "n": n, "t_end": t_end, "blocks": b.blocks, "particle_steps": b.particle_steps,
            "seconds": round(wall, 3), "seconds_in_abi": round(b.abi_seconds, 3),
            "phase_seconds": {k: round(v, 3) for k, v in b.phase_seconds.items()},
            "pairs_per_s": b.particle_steps * (n - 1) / wall,
            "de_over_e": abs((e1 - e0) / e0)}, c


def kira_callmix_leg(c, t_end=2.0 ** -8):
    """Config 3's cluster with kira's whole GRAPE-6 call mix (bs_run_kira):
    perturber lists and a perturbation pair call for every binary centre of
    mass at every step, list overflows and single-node redos as kira makes
    them; timed end to end, no checks."""
    from starlab_amd.blockstep import KiraCallMix
    n = len(c["m"])
    bins = dict(node=c["bin_index"], m1=c["m1"], dx=c["dx"], dv=c["dv"])
    k = KiraCallMix(c["m"], c["x"], c["v"], eta=0.1, eps2=0.0).run(t_end, binaries=bins)
    st = dict(k.stats)
    return {"workload": f"King W0=6 N={n}, {len(bins['node'])} binaries with their orbits, "
                        f"t=0->{t_end}; bs_run_kira (lists, perturbation pairs, redos)",
            "seconds": round(st.pop("seconds"), 3),
            "seconds_in_abi": round(st.pop("abi_seconds"), 3),
            "pairs_per_s_blocks": k.particle_steps * (n - 1) / k.seconds, "calls": st}


TRAFFIC_FILE = os.path.join("profiles", "force_n65536_traffic.json")


def load_traffic():
    """HBM bytes per k_force launch of this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in TRAFFIC_FILE (counters cannot
    run inside this timed process); None if absent."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch"), d.get("measured")
    except (OSError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ens", action="store_true", help="skip the ensemble / serving legs")
    ap.add_argument("--no-kira", action="store_true",
                    help="skip the kira end-to-end leg (native block-step driver)")
    ap.add_argument("--no-tune", action="store_true", help="skip the force-shape autotuner")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the energy/density fast-path and GRAPE-4 legs")
    ap.add_argument("--tune-file", default=None,
                    help="force-shape table: loaded if it exists (no tuning), else written "
                         "after tuning")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 256 MiB streaming kernel (read + write), the "
                         "reference for scaling HBM traffic counters, and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "libg6 drives one GPU per process; this benchmark measures "
                                   "N=1 only (multi-GPU sharding is a design note)"}))
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    if a.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy": "torch.neg float32", "bytes_read": src.numel() * 4,
                          "bytes_written": dst.numel() * 4, "launches": 5}))
        return

    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    if g.open() != 0:
        raise SystemExit(f"g6_open_ failed: {g.last_error()}")
    n = a.n
    tune = None
    if a.tune_file and os.path.exists(a.tune_file):
        nb = g.load_tuning(a.tune_file)
        tune = {"loaded": a.tune_file, "buckets": nb}
    log(f"bench: generating Plummer N={n}")
    inp = make_input(g, n)

    if tune is None and not a.no_tune:
        log("bench: autotuning force shapes")
        t0 = time.perf_counter()
        nb = g.autotune(n, n, EPS2)
        tune = {"seconds": round(time.perf_counter() - t0, 3), "buckets": nb}
        if a.tune_file:
            g.save_tuning(a.tune_file)
    if tune is not None:
        full = g.tuning().get(max(0, (n - 1).bit_length()))
        tune["full_sweep_shape"] = (dict(zip(("variant", "tile", "waves", "lanes"), full))
                                    if full else None)

    for _ in range(a.warmup):
        assert g.self_sweep(n, 0, n, 0.0, EPS2) == 0
    g.sync()
    g.option("timing", 1)
    g.reset_timing()
    g.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        assert g.self_sweep(n, 0, n, 0.0, EPS2) == 0
    g.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    g.option("timing", 0)

    pairs_per_step = n * (n - 1)
    value = pairs_per_step * a.steps / elapsed
    kcalls, kms = g.kernel_time("force")
    force_ms = kms / max(kcalls, 1)
    achieved = FLOP_PER_PAIR * pairs_per_step / (force_ms * 1e-3) / 1e12
    breakdown = {}
    for k in ("predict", "force", "reduce"):
        c, ms = g.kernel_time(k)
        breakdown[k] = round(ms / max(c, 1), 4)
    wg, nsplit, jchunk = g.geometry()
    acc, jerk, pot, nn = g.read_results(n)
    refinements = g.refinements()
    traffic, traffic_when = load_traffic()
    # config 2's results and j-memory, kept for the error against FP64
    gpu = {"acc": acc, "jerk": jerk, "pot": pot, "nn": nn, "samples": [],
           "jm": {"t": inp["tj"], "m": inp["m"], "x": inp["x"], "v": inp["v"],
                  "a2": inp["a2"], "j6": inp["j6"], "k18": np.zeros((n, 3))}}

    out = {
        "metric": "pair interactions/s at N=64K vs FP32 roofline; ensemble p50/p99 latency",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "force-only: Plummer N=65536, equal masses, N-body units, "
                               "eps2=1e-8, all i vs all j, t_j spread U[0,2^-10], "
                               "double-single positions, FP32 pair math, FP64 accumulation",
                   "n": n, "eps2": EPS2, "pairs_per_step": pairs_per_step},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "traffic_source": f"{TRAFFIC_FILE} ({traffic_when})",
                     "kernel": "k_force", "kernel_ms": force_ms,
                     "flop_per_pair": FLOP_PER_PAIR},
        "kernel_ms": breakdown,
        "force_geometry": {"workgroups": wg, "nsplit": nsplit, "jchunk": jchunk,
                           "tile": g.option("last_tile")},
        "nn_refinements": refinements,
        "autotune": tune,
    }
    if not a.no_aux:
        log("bench: energy / density fast paths")
        out["aux"] = aux_legs(g, inp)
        assert g.self_sweep(n, 0, n, 0.0, EPS2) == 0
        xp, vp = g.read_predicted(n)
    g.close()
    if not a.no_aux:
        log("bench: GRAPE-4 ABI sweep")
        out["aux"]["grape4_abi_sweep"] = grape4_leg(inp, xp, vp)
        log("bench: GRAPE-6 ABI block latency (C caller)")
        out["aux"]["grape6_abi_blocks"] = abi_leg()
        # about 30 neighbours per i in the uniform sphere of abi_bench
        out["aux"]["grape6_abi_blocks_with_lists"] = abi_leg(blocks=100, sizes=(1, 256, 2048),
                                                             h2=0.006)
    kira = None
    if not a.no_kira:
        log("bench: config 1 (native block-step driver, Plummer N=1024 to t=1)")
        s1 = []
        out["config1"] = config1_leg(s1)
        gpu["samples"].append(("config1", s1, 0.0))
        log("bench: kira end to end (native block-step driver, N=262144 to t=1/4)")
        s3 = []
        kira, cluster = kira_leg(samples=s3)
        out["kira_end_to_end"] = kira
        gpu["samples"].append(("config3", s3, 0.0))
        log("bench: kira's call mix on config 3 (bs_run_kira, to t=1/256)")
        out["kira_callmix"] = kira_callmix_leg(cluster)
        del cluster
    clusters = None
    if not a.no_ens:
        log("bench: ensemble and serving legs")
        ens = ensemble_legs()
        out["ensemble"] = ens
        out["ensemble_p50_s"] = ens["serving"]["p50_s"]
        out["ensemble_p99_s"] = ens["serving"]["p99_s"]
        from starlab_amd import serve
        stride = 64
        clusters = {"sample": serve.cluster_batch(4096, 32, 2048, 4, 1.0 / 32)[::stride],
                    "stride": stride,
                    "gpu_pairs_per_s": ens["batch_fp32_t1"]["pairs_per_s"],
                    "gpu_pairs_t1": ens["batch_fp32_t1"]["pairs"]}
    if not a.no_cpu:
        log("bench: CPU baselines and errors against FP64")
        cb, errs = cpu_leg(inp, gpu, clusters=clusters, kira=kira)
        out["cpu_baseline"] = cb
        if "ensemble" in out:
            e = out["ensemble"]["error_vs_fp64"]
            errs["config4"] = e
            errs["config5"] = {"3body": "FP64 pairs on block steps: the block-step oracle "
                                        "integrator bit for bit (tests/test_gpu_ens_block.py)",
                               "clusters": "FP32 pairs, as config4"}
        out["error_vs_fp64"] = errs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
