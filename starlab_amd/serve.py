"""Ensemble workloads on the batched engine: a fixed batch (all jobs at once)
and a served Poisson stream (continuous batching, p50/p99 latency).

Jobs are generated before the clock starts; latency is measured from each
job's scheduled arrival to the moment the host sees it finished, so queueing
for slots / HBM budget is included.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import load
from .ens import Ensemble, autotune, binary_single, plummer_cluster

GB = 1 << 30
# the engine's step kernels, as named by its kernel_us_<name> counters
KERNELS = ("predict", "force", "correct", "bselect", "bitems", "bforce", "bcorrect", "finalize",
           "fused")


def cluster_job(n, rng, t_end, eta=0.1, prec=32):
    m, x, v = plummer_cluster(n, rng)
    return dict(kind="cluster", n=n, m=m, x=x, v=v, t_end=t_end, eta=eta, eps2=(4.0 / n) ** 2,
                prec=prec)


def scatter_job(rng, t_end, eta=0.1, r_init=10.0, prec=64):
    m, x, v = binary_single(rng, r_init=r_init)
    return dict(kind="3body", n=3, m=m, x=x, v=v, t_end=t_end, eta=eta, eps2=0.0, prec=prec)


def cluster_batch(count, nmin, nmax, seed, t_end, eta=0.1, prec=32):
    """`count` Plummer clusters with N uniform in [nmin, nmax]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, count)
    return [cluster_job(int(n), rng, t_end, eta, prec) for n in sizes]


def _submit(eng, j):
    eng.option("precision", j.get("prec", 64))
    return eng.submit(j["m"], j["x"], j["v"], j["t_end"], j["eta"], j["eps2"])


def serving_mix(count, seed, frac_3body=0.7, nmin=256, nmax=4096, t3=32.0, tc=1.0 / 32,
                eta=0.1):
    """The serving workload: 70 % binary-single scatterings (FP64 pairs), 30 %
    clusters (FP32 pairs)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        if rng.uniform() < frac_3body:
            jobs.append(scatter_job(rng, t3, eta))
        else:
            jobs.append(cluster_job(int(rng.integers(nmin, nmax + 1)), rng, tc, eta))
    return jobs


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def run_batch(jobs, budget=64 * GB, max_systems=8192, chunk=32, tune=False, keep=False, **opt):
    """Submit every job at t = 0 and run to completion.  With tune=True the
    engine's shape options come from ens.autotune on a sample of the jobs
    (its time is reported, not counted in wall_s); keep=True adds each job's
    result (x, v, t, steps, e0, e1, failed) in job order.

    Block-step jobs default to two heavy lanes (results are bitwise the same
    as with one): config 4's 4096 clusters to t = 1 took 2.52-2.73 s against
    2.97-3.03 s on one."""
    block = opt.get("block_steps")
    if block is None:
        block = load().g6e_get_option(b"block_steps")
    if block:
        opt.setdefault("heavy_lanes", 2)
    tuned = None
    if tune:
        # the tuner's candidates run with the step mode and lanes the batch will use
        mode = {k: v for k, v in opt.items() if k in ("block_steps", "block_fusion", "heavy_lanes")}
        best, timings, secs = autotune(jobs, max_systems=max_systems, fixed=mode)
        opt = {**opt, **best}
        chunk = best.get("steps_per_graph", chunk)
        tuned = dict(options=best, seconds=round(secs, 3), candidates=len(timings))
    eng = Ensemble(budget, max_systems, **{"steps_per_graph": chunk, **opt})
    try:
        t0 = time.perf_counter()
        ids = [_submit(eng, j) for j in jobs]
        done_at = {}
        results = {}
        while True:
            left = eng.run(chunk)
            now = time.perf_counter() - t0
            for jid in eng.poll():
                done_at[jid] = now
                results[jid] = eng.result(jid)
            if left == 0:
                break
        wall = time.perf_counter() - t0
        st = eng.stats()
        kernel_us = None
        if opt.get("kernel_timing"):
            kernel_us = {k: eng.option("kernel_us_" + k) for k in KERNELS}
            kernel_us = {k: v for k, v in kernel_us.items() if v}
    finally:
        eng.close()
    lat = [done_at[i] for i in ids]
    de = [abs((results[i]["e1"] - results[i]["e0"]) / results[i]["e0"]) for i in ids]
    out = dict(jobs=len(jobs), wall_s=wall, systems_per_s=len(jobs) / wall,
               pairs=st["pairs"], pairs_per_s=st["pairs"] / wall,
               device_s=st["device_seconds"], engine_steps=st["engine_steps"],
               p50_s=_pct(lat, 50), p99_s=_pct(lat, 99),
               failed=sum(results[i]["failed"] for i in ids),
               de_median=_pct(de, 50), de_max=float(np.max(de)), tuned=tuned)
    if kernel_us is not None:
        out["kernel_us"] = kernel_us
    if keep:
        out["results"] = [results[i] for i in ids]
    return out


def run_poisson(jobs, rate, seed=0, budget=64 * GB, max_systems=4096, chunk=16, keep=False,
                **opt):
    """Serve `jobs` arriving as a Poisson stream of `rate` per second;
    keep=True adds each job's result in job order.

    Each engine call runs `chunk` steps of the heavy lane (one graph) while
    the light lane's fused kernel takes up to `fused_steps` steps per launch
    (default here 64).  Arrivals wait for the current call, so longer calls
    raise the small jobs' latency and shorten the clusters'; on the serving
    mix at 2000 req/s, 16 and 64 measured p99 37 ms with p50 6.1 ms, against
    p99 57 ms at 8 and 256, and p50 8.3 ms at 64 and 64."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    opt.setdefault("fused_steps", 64)
    eng = Ensemble(budget, max_systems, steps_per_graph=chunk, **opt)
    try:
        t0 = time.perf_counter()
        nxt = 0
        arrived = {}
        done_at = {}
        results = {}
        kind = {}
        order = []
        while len(done_at) < len(jobs):
            now = time.perf_counter() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                j = jobs[nxt]
                jid = _submit(eng, j)
                order.append(jid)
                arrived[jid] = arrivals[nxt]
                kind[jid] = j["kind"]
                nxt += 1
            left = eng.run(chunk)
            now = time.perf_counter() - t0
            for jid in eng.poll():
                done_at[jid] = now
                results[jid] = eng.result(jid)
            if left == 0 and nxt < len(jobs):
                wait = arrivals[nxt] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
        wall = time.perf_counter() - t0
        st = eng.stats()
    finally:
        eng.close()
    lat = {k: [] for k in ("3body", "cluster")}
    allat = []
    for jid, t in done_at.items():
        d = t - arrived[jid]
        lat[kind[jid]].append(d)
        allat.append(d)
    out = dict(jobs=len(jobs), rate_req_s=rate, offered_s=float(arrivals[-1]), wall_s=wall,
               throughput_req_s=len(jobs) / wall, pairs_per_s=st["pairs"] / wall,
               device_busy=st["device_seconds"] / wall,
               p50_s=_pct(allat, 50), p99_s=_pct(allat, 99),
               p50_3body_s=_pct(lat["3body"], 50), p99_3body_s=_pct(lat["3body"], 99),
               p50_cluster_s=_pct(lat["cluster"], 50), p99_cluster_s=_pct(lat["cluster"], 99),
               failed=sum(r["failed"] for r in results.values()))
    if keep:
        out["results"] = [results[j] for j in order]
    return out


def main(argv=None):
    """Command line: run the batched-cluster workload or the served Poisson
    stream on the ensemble engine and print one JSON line.

        python -m starlab_amd.serve batch [--count 4096 --nmin 32 --nmax 2048
                                           --t-end 0.03125 --prec 32 --tune
                                           --block-steps 1 --heavy-lanes 2
                                           --kernel-timing]
        python -m starlab_amd.serve poisson [--rate 2000 --seconds 2]
    """
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m starlab_amd.serve")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("batch", help="Plummer clusters, N ~ U[nmin, nmax], all at t = 0")
    b.add_argument("--count", type=int, default=4096)
    b.add_argument("--nmin", type=int, default=32)
    b.add_argument("--nmax", type=int, default=2048)
    b.add_argument("--t-end", type=float, default=1.0 / 32)
    b.add_argument("--prec", type=int, choices=(32, 64), default=32)
    b.add_argument("--seed", type=int, default=4)
    b.add_argument("--tune", action="store_true", help="engine shapes from ens.autotune")
    b.add_argument("--block-steps", type=int, choices=(0, 1), default=1,
                   help="1: kira's individual block steps; 0: one shared step per system")
    b.add_argument("--heavy-lanes", type=int, choices=(1, 2), default=None,
                   help="heavy systems on one or two streams (default: 2 for block steps)")
    b.add_argument("--kernel-timing", action="store_true",
                   help="step kernels one at a time between HIP events; adds kernel_us")
    p = sub.add_parser("poisson", help="70 %% three-body scatterings, 30 %% clusters")
    p.add_argument("--rate", type=float, default=2000.0)
    p.add_argument("--seconds", type=float, default=2.0)
    p.add_argument("--seed", type=int, default=5)
    a = ap.parse_args(argv)
    if a.cmd == "batch":
        jobs = cluster_batch(a.count, a.nmin, a.nmax, a.seed, a.t_end, prec=a.prec)
        opt = {"block_steps": a.block_steps}
        if a.kernel_timing:
            opt["kernel_timing"] = 1
        if a.heavy_lanes is not None:
            opt["heavy_lanes"] = a.heavy_lanes
        r = run_batch(jobs, tune=a.tune, **opt)
    else:
        r = run_poisson(serving_mix(int(a.rate * a.seconds), a.seed), a.rate)
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
