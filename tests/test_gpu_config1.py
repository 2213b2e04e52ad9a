"""Config 1: plain kira-style integration of a Plummer sphere (N = 1024,
seed 42 as in the reference's init Testfile, equal masses, N-body units,
eps = 0, eta = 0.1) from t = 0 to t = 1 with individual power-of-two block
steps, forces from libg6 through the GRAPE-6 ABI.

* every block step: libg6's acc/jerk/pot agree with the FP64 oracle on the
  same inputs and the nearest-neighbour index is exact.  Bars, for every
  particle step: normwise acc and jerk error (|da| / sum_j |a_ij|) <= 1e-6
  and pot <= 1e-6; per-particle acc error |da| / |a| <= 1e-6 for 99.9 % of
  particle steps (the rest are steps whose pair terms cancel to |a| well
  below sum_j |a_ij|, where FP32 pair rounding is amplified; measured max
  2.4e-6 in one of 257725 steps, normwise 1.5e-7).  Per-particle
  percentiles (median, p99, max) of acc, jerk and pot are printed;
* the energy error |dE/E| at t = 1 stays within a factor 10 (+1e-8) of the
  all-CPU FP64 run (kira's host force path) from the same initial state.
"""
import time

import numpy as np
import pytest

import oracle
from starlab_amd.blockstep import BlockHermite
from tests.kira_backends import CheckedG6, HostBackend

pytestmark = pytest.mark.gpu


def _energy(b):
    return b.energy(lambda m, x, v, e: oracle.energies(m, x, v, e))


def test_config1_plummer_1024_to_t1():
    from starlab_amd.g6 import GRAPE6
    m, x, v = oracle.makeplummer(1024, 42)

    cpu = BlockHermite(m, x, v, eta=0.1, eps2=0.0)
    e0 = _energy(cpu)
    t0 = time.perf_counter()
    cpu.run(HostBackend(), 1.0)
    t_cpu = time.perf_counter() - t0
    de_cpu = abs((_energy(cpu) - e0) / e0)

    gpu = BlockHermite(m, x, v, eta=0.1, eps2=0.0)
    be = CheckedG6(GRAPE6())
    t0 = time.perf_counter()
    gpu.run(be, 1.0)
    t_gpu = time.perf_counter() - t0
    de_gpu = abs((_energy(gpu) - e0) / e0)
    print(f"cpu: blocks {cpu.blocks} steps {cpu.particle_steps} dE/E {de_cpu:.3e} {t_cpu:.1f}s",
          flush=True)
    print(f"gpu: blocks {gpu.blocks} steps {gpu.particle_steps} dE/E {de_gpu:.3e} {t_gpu:.1f}s "
          f"checked {be.checked_blocks} worst {be.worst} nn_mismatch {be.nn_mismatch}", flush=True)
    assert np.all(gpu.t == 1.0)
    assert be.checked_blocks == gpu.blocks + 1
    assert be.nn_mismatch == 0
    print(f"errors vs FP64: {be.summary()}", flush=True)
    assert be.worst["acc_norm"] <= 1e-6
    assert be.worst["jerk_norm"] <= 1e-6
    assert np.percentile(np.concatenate(be.rel_errs["acc"]), 99.9) <= 1e-6
    assert be.worst["pot_rel"] <= 1e-6
    assert de_gpu <= 10 * de_cpu + 1e-8


def test_config1_native_driver_equals_python_driver():
    """The native block-step driver (libkira_blockstep.so) on libg6 reproduces
    the Python driver on libg6 bit for bit -- per-particle j updates as kira
    makes them, and the bulk j update -- for config 1 to t = 1/4; the
    observer's oracle checks hold on every block."""
    from starlab_amd.blockstep import G6Backend, NativeBlockHermite
    from starlab_amd.g6 import GRAPE6
    from tests.kira_backends import NativeChecker
    m, x, v = oracle.makeplummer(1024, 42)
    py = BlockHermite(m, x, v, eta=0.1, eps2=0.0)
    py.run(G6Backend(GRAPE6()), 0.25)
    chk = NativeChecker(0.0)
    nat = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(0.25, observer=chk)
    bulk = NativeBlockHermite(m, x, v, eta=0.1, eps2=0.0).run(0.25, bulk=True)
    print(f"python {py.blocks} blocks; native {nat.blocks} blocks {nat.seconds:.2f}s "
          f"(ABI {nat.abi_seconds:.2f}s); bulk {bulk.seconds:.2f}s", flush=True)
    for other in (nat, bulk):
        assert other.blocks == py.blocks and other.particle_steps == py.particle_steps
        for name in ("x", "v", "a", "j", "pot", "dt"):
            np.testing.assert_array_equal(getattr(other, name), getattr(py, name), err_msg=name)
    assert chk.checked_blocks == nat.blocks and chk.nn_mismatch == 0
    assert chk.worst["acc_norm"] <= 1e-6 and chk.worst["jerk_norm"] <= 1e-6
    assert np.percentile(np.concatenate(chk.rel_errs["acc"]), 99.9) <= 1e-6
    assert chk.worst["pot_rel"] <= 1e-6


def test_config1_snapshot_restart():
    """kira's restart (-R, kira.C:46-113; the j-memory is reloaded after
    g6_open_, hdyn_grape6.C:3004-3054): config 1's initial conditions go
    through a Starlab snapshot (bit-exact round trip), the run stops at
    t = 1/8, writes a snapshot, closes libg6, and a new run restarts from the
    snapshot read back to t = 1/4.  Every restarted block agrees with the
    FP64 oracle, the restart state is the stopped state bit for bit, and the
    energy change over the two legs stays within config 1's factor of the
    all-CPU run."""
    import io

    from starlab_amd import snapshot
    from starlab_amd.blockstep import NativeBlockHermite
    from tests.kira_backends import NativeChecker
    m, x, v = oracle.makeplummer(1024, 42)
    root = snapshot.read_snapshot(io.StringIO(snapshot.write_snapshot(None,
                                                                      snapshot.from_arrays(m, x, v))))
    m0, x0, v0 = snapshot.top_level(root)
    for a, b in ((m0, m), (x0, x), (v0, v)):
        np.testing.assert_array_equal(a, b)
    e0 = sum(oracle.energies(m, x, v, 0.0))
    first = NativeBlockHermite(m0, x0, v0, eta=0.1, eps2=0.0).run(0.125)
    text = snapshot.write_snapshot(None, snapshot.from_arrays(first.m, first.x, first.v,
                                                              system_time=0.125))
    back = snapshot.read_snapshot(io.StringIO(text))
    assert back.system_time == 0.125
    m1, x1, v1 = snapshot.top_level(back)
    np.testing.assert_array_equal(x1, first.x)
    np.testing.assert_array_equal(v1, first.v)
    chk = NativeChecker(0.0)
    second = NativeBlockHermite(m1, x1, v1, eta=0.1, eps2=0.0).run(0.125, observer=chk)
    cpu = BlockHermite(m, x, v, eta=0.1, eps2=0.0)
    cpu.run(HostBackend(), 0.25)
    de_cpu = abs((_energy(cpu) - e0) / e0)
    de = abs(sum(oracle.energies(second.m, second.x, second.v, 0.0)) / e0 - 1)
    print(f"restart: {first.blocks} + {second.blocks} blocks, dE/E {de:.3e} (all-CPU, "
          f"uninterrupted: {de_cpu:.3e}); {chk.summary()}", flush=True)
    assert chk.checked_blocks == second.blocks and chk.nn_mismatch == 0
    assert chk.worst["acc_norm"] <= 1e-6 and chk.worst["jerk_norm"] <= 1e-6
    assert chk.worst["pot_rel"] <= 1e-6
    assert de <= 10 * de_cpu + 1e-8
