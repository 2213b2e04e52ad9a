"""Config 3: King W0 = 6, N = 262144 top-level nodes, Salpeter masses
(slope -2.35, mass ratio 100), 5 % primordial binaries (log-uniform a,
thermal e) as their centres of mass -- what kira keeps in GRAPE j-memory --
integrated with kira's individual power-of-two block steps from t = 0 to
t = 0.25 by the native block-step driver (libkira_blockstep.so), forces from
libg6 through the GRAPE-6 ABI (eps = 0, eta = 0.1).

* a seeded sample of block steps (every 97th force call, at most 16 i each)
  is recomputed by the FP64 oracle on the same inputs: acc and jerk
  normwise error <= 1e-6 and pot <= 1e-6 for every sampled particle step,
  per-particle acc error |da|/|a| <= 1e-6 for 99.9 % of them (the 100:1
  masses let a heavy j that is not the nearest dominate, and sums that
  cancel 5-20 fold amplify the FP32 pair rounding), nearest-neighbour index
  exact; per-particle percentiles of acc, jerk and pot are printed;
* every particle lands on t = 0.25 and block times are exact binary
  fractions (the power-of-two schedule);
* the total energy (FP64 oracle pair sum) changes by |dE/E| <= 1e-4.

The same model at N = 4096, a size the CPU integrates in seconds, runs on
libg6 and on the FP64 oracle behind the same ABI: libg6's energy drift stays
within 10x the all-CPU run's (+ 1e-8).
"""
import os
import subprocess
import time

import numpy as np
import pytest

import oracle
from starlab_amd import ics
from starlab_amd.blockstep import NativeBlockHermite
from tests.kira_backends import NativeChecker

pytestmark = pytest.mark.gpu

N = 262144
T_END = 0.25
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_config3_king_salpeter_binaries_262144_to_quarter():
    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    assert g.open() == 0, g.last_error()
    idx = np.arange(N, dtype=np.int32)
    z = np.zeros((N, 3))

    def dev_pot(m, x):
        g.set_j_particles(idx, idx, np.zeros(N), np.ones(N), m, z, z, z, z, x)
        assert g.self_sweep(N, 0, N, 0.0, 0.0) == 0
        return 0.5 * float((m * g.read_results(N)[2]).sum())

    t0 = time.perf_counter()
    c = ics.king_cluster(N, 42, potential_fn=dev_pot)
    g.close()
    m, x, v = c["m"], c["x"], c["v"]
    print(f"ics {time.perf_counter() - t0:.1f}s binaries {len(c['bin_index'])} "
          f"m_max/m_min {m.max() / m.min():.1f}", flush=True)
    e0 = sum(oracle.energies(m, x, v, 0.0))

    chk = NativeChecker(0.0, every=97, max_i=16, seed=3)
    sizes, times = [], []
    obs_s = [0.0]
    last = [time.perf_counter()]

    def observer(jm, t, idx, xi, vi, acc, jerk, pot, nn):
        t0 = time.perf_counter()
        sizes.append(len(idx))
        times.append(t)
        chk(jm, t, idx, xi, vi, acc, jerk, pot, nn)
        now = time.perf_counter()
        obs_s[0] += now - t0
        if now - last[0] > 10:
            last[0] = now
            print(f"  t={t:.5f} blocks {len(sizes)}", flush=True)
        return False

    # kira's per-block sequence in native code (libkira_blockstep.so) on libg6
    b = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(T_END, observer=observer)
    wall = b.seconds - obs_s[0]
    be = chk
    e1 = sum(oracle.energies(b.m, b.x, b.v, 0.0))
    de = abs((e1 - e0) / e0)
    errs = np.concatenate(be.acc_errs)
    sizes = np.array(sizes)
    hist = {int(1 << k): int(((sizes >= (1 << k)) & (sizes < (2 << k))).sum())
            for k in range(0, 19)}
    print(f"blocks {b.blocks} particle steps {b.particle_steps} wall {wall:.1f}s without the "
          f"checks (in ABI calls {b.abi_seconds:.1f}s, checks {obs_s[0]:.1f}s) "
          f"pair interactions/s {b.particle_steps * (N - 1) / wall:.3e} dE/E {de:.3e}", flush=True)
    print(f"block sizes (lower edge: count) {hist}", flush=True)
    print(f"checked {be.checked_blocks} calls / {be.checked_i} particle steps; "
          f"acc p99.9 {np.percentile(errs, 99.9):.2e} max {errs.max():.2e} worst {be.worst} "
          f"nn_mismatch {be.nn_mismatch}", flush=True)
    assert np.all(b.t == T_END)
    assert all(abs(t * 2.0 ** 40 - round(t * 2.0 ** 40)) == 0 for t in times)
    assert np.all(np.log2(b.dt) == np.round(np.log2(b.dt)))
    assert be.checked_blocks >= 400 and be.nn_mismatch == 0
    summ = be.summary()
    print(f"errors vs FP64: {summ}", flush=True)
    assert be.worst["acc_norm"] <= 1e-6
    assert be.worst["jerk_norm"] <= 1e-6
    assert be.worst["pot_rel"] <= 1e-6
    assert np.percentile(np.concatenate(be.rel_errs["acc"]), 99.9) <= 1e-6
    assert de <= 1e-4


def test_config3_model_energy_drift_against_cpu_run(tmp_path):
    """Config 3's model -- King W0 = 6, Salpeter 1-100, 5 % binaries as their
    centres of mass, the same generator and seed -- at N = 4096, integrated
    to t = 0.25 twice by the native driver: forces from libg6, and from the
    FP64 oracle behind the same GRAPE-6 ABI on the host (tests/oracle_g6abi.c,
    bit-exact oracle build).  A CPU run at N = 262144 takes ~2400 s of force
    alone, so the drift comparison runs at this size, the full-size run
    above keeps its absolute bound.  libg6's |dE/E| stays within 10x the
    all-CPU run's + 1e-8, the bound of config 1; both runs land every
    particle on t = 0.25."""
    lib = str(tmp_path / "liboracle_g6abi.so")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'oracle')}",
                    os.path.join(ROOT, "tests", "oracle_g6abi.c"),
                    os.path.join(ROOT, "oracle", "g6_oracle.c"), "-lm", "-o", lib], check=True)
    n = 4096
    c = ics.king_cluster(n, 42)
    m, x, v = c["m"], c["x"], c["v"]
    e0 = sum(oracle.energies(m, x, v, 0.0))
    gpu = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(T_END)
    os.environ["ORACLE_G6ABI_JSPLIT"] = "16"
    try:
        cpu = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(T_END, grape_lib=lib)
    finally:
        del os.environ["ORACLE_G6ABI_JSPLIT"]
    de_gpu = abs(sum(oracle.energies(gpu.m, gpu.x, gpu.v, 0.0)) / e0 - 1)
    de_cpu = abs(sum(oracle.energies(cpu.m, cpu.x, cpu.v, 0.0)) / e0 - 1)
    dx = np.linalg.norm(gpu.x - cpu.x, axis=1)
    print(f"N={n} binaries {len(c['bin_index'])}: gpu blocks {gpu.blocks} steps "
          f"{gpu.particle_steps} {gpu.seconds:.2f}s dE/E {de_gpu:.3e}; cpu blocks {cpu.blocks} "
          f"steps {cpu.particle_steps} {cpu.seconds:.2f}s dE/E {de_cpu:.3e}; |x_gpu - x_cpu| "
          f"median {np.median(dx):.2e} max {dx.max():.2e}", flush=True)
    assert np.all(gpu.t == T_END) and np.all(cpu.t == T_END)
    assert de_gpu <= 10 * de_cpu + 1e-8
