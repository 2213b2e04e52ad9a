"""kira's GRAPE-6 call mix (bs_run_kira, libkira_blockstep.so) on libg6:
neighbour radii and list rounds, single-node redos, perturbation pairs of
binaries, energy sweeps with the reload after them, density rounds, release
and re-attach, and reset-and-retry after host-injected errors, each made as
kira's driver makes it (hdyn_grape6.C).  Every checked force call of every
kind agrees with the FP64 oracle on the j-memory exactly as sent (normwise
acc / jerk and pot <= 1e-6, nearest neighbour exact); every checked
neighbour list equals the brute-force FP64 set (i itself included, overflow
iff 4096 or more entries)."""
import os
import subprocess
import time

import numpy as np
import pytest

import oracle
from starlab_amd.blockstep import (CALL_BLOCK, CALL_DENSITY, CALL_ENERGY, CALL_LIST,
                                   CALL_PERTURB, CALL_REDO, KiraCallMix)
from tests.kira_backends import CallMixChecker
from tests.test_kira_callmix_cpu import binaries

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _check_bars(chk):
    print(f"checked {chk.checked_blocks} calls / {chk.checked_i} i, worst {chk.worst}, "
          f"nn_mismatch {chk.nn_mismatch}, lists {chk.list_checked} "
          f"(mismatches {chk.list_mismatch[:3]}), kinds {dict(chk.kinds)}, "
          f"reads {dict(chk.read_status)}, summary {chk.summary()}", flush=True)
    assert chk.nn_mismatch == 0 and chk.list_mismatch == []
    assert chk.worst["acc_norm"] <= 1e-6 and chk.worst["jerk_norm"] <= 1e-6
    assert chk.worst["pot_rel"] <= 1e-6


def _oracle_shim(tmp_path):
    out = tmp_path / "liboracle_g6abi.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'oracle')}",
                    os.path.join(ROOT, "tests", "oracle_g6abi.c"),
                    os.path.join(ROOT, "oracle", "g6_oracle.c"), "-lm", "-o", str(out)],
                   check=True)
    return str(out)


def test_callmix_plummer_1024_every_call(tmp_path, monkeypatch):
    """Config 1's system (Plummer N = 1024, seed 42, eps = 0, eta = 0.1) with
    stellar radii, 16 binary centres of mass, energies and densities every
    1/8, a release every 400 blocks, to t = 1/4.  The same rules on the FP64
    oracle library give the CPU run; the energy change of the two agrees
    within a factor 10 (+1e-8)."""
    m, x, v = oracle.makeplummer(1024, 42)
    kw = dict(star_radius=0.005, binaries=binaries(m, 2, 16), dt_log=0.125, densities=True,
              release_every=400)
    e0 = sum(oracle.energies(m, x, v, 0.0))
    chk = CallMixChecker(0.0, npipes=2048)
    t0 = time.perf_counter()
    g = KiraCallMix(m, x, v).run(0.25, observer=chk, **kw)
    wall = time.perf_counter() - t0
    monkeypatch.setenv("ORACLE_G6ABI_NPIPES", "2048")
    c = KiraCallMix(m, x, v).run(0.25, grape_lib=_oracle_shim(tmp_path), **kw)
    de_g = abs(sum(oracle.energies(g.m, g.x, g.v, 0.0)) / e0 - 1)
    de_c = abs(sum(oracle.energies(c.m, c.x, c.v, 0.0)) / e0 - 1)
    print(f"gpu {g.stats} {wall:.1f}s dE/E {de_g:.3e}; cpu dE/E {de_c:.3e} "
          f"blocks {c.stats['blocks']}", flush=True)
    _check_bars(chk)
    assert chk.checked_blocks == g.stats["force_calls"]
    assert chk.list_checked == g.stats["list_gets"] > 1000
    for kind in (CALL_BLOCK, CALL_REDO, CALL_PERTURB, CALL_ENERGY, CALL_DENSITY, CALL_LIST):
        assert chk.kinds[kind] > 0, kind
    s = g.stats
    assert s["energy_sweeps"] == s["density_rounds"] == 3 and s["reopens"] > 0
    assert np.all(g.t == 0.25)
    assert de_g <= 10 * de_c + 1e-8
    np.testing.assert_allclose(g.energies, c.energies, rtol=1e-4)


def test_callmix_recovery_on_libg6():
    """Errors injected on the host side of libg6 (options fail_lasthalf,
    fail_read_nl, fail_nn; the device work runs normally) make the driver
    reset, reload and retry as kira does: bitwise the clean run."""
    from starlab_amd import load
    L = load()
    m, x, v = oracle.makeplummer(512, 7)
    kw = dict(star_radius=0.01, dt_log=1.0 / 32)
    clean = KiraCallMix(m, x, v).run(0.0625, **kw)
    for k, n in ((b"fail_lasthalf", 3), (b"fail_read_nl", 2), (b"fail_nn", 1)):
        assert L.g6x_set_option(k, n) == 0
    try:
        chk = CallMixChecker(0.0, npipes=2048)
        bad = KiraCallMix(m, x, v).run(0.0625, observer=chk, **kw)
    finally:
        for k in (b"fail_lasthalf", b"fail_read_nl", b"fail_nn"):
            L.g6x_set_option(k, 0)
    print(bad.stats, flush=True)
    assert clean.stats["resets"] == 0
    assert bad.stats["resets"] == 6 and bad.stats["error42"] == 1
    assert chk.errors == 4 and chk.read_status[-1] == 2
    _check_bars(chk)
    for name in ("x", "v", "a", "j", "pot", "dt"):
        np.testing.assert_array_equal(getattr(bad, name), getattr(clean, name), err_msg=name)
    np.testing.assert_array_equal(bad.energies, clean.energies)


def test_callmix_list_overflows_on_libg6():
    """Both overflow paths at full list size: lists of 4096+ entries
    (get_neighbour_list overflow: radius x 0.25, single-node redo), and a
    small list store (option nb_cap) that makes read_neighbour_list report
    overflow (radii halved, chunk redone)."""
    from starlab_amd import load
    L = load()
    m, x, v = oracle.makeplummer(6000, 3)
    chk = CallMixChecker(0.0, npipes=2048, every=97, max_i=8, list_every=53)
    k = KiraCallMix(m, x, v).run(2.0 ** -10, star_radius=3.0, observer=chk)
    print(k.stats, flush=True)
    assert k.stats["get_overflows"] > 1000 and k.stats["redo_calls"] >= k.stats["get_overflows"]
    _check_bars(chk)
    assert chk.list_checked > 100
    old = L.g6x_get_option(b"nb_cap")
    assert L.g6x_set_option(b"nb_cap", 24) == 0
    try:
        chk = CallMixChecker(0.0, npipes=2048, list_every=7)
        m, x, v = oracle.makeplummer(512, 9)
        k = KiraCallMix(m, x, v).run(2.0 ** -6, star_radius=0.3, observer=chk)
    finally:
        L.g6x_set_option(b"nb_cap", old)
    print(k.stats, flush=True)
    assert k.stats["read_overflows"] > 0
    _check_bars(chk)


def test_config3_callmix_sample():
    """Config 3's cluster (King W0 = 6, N = 262144 top-level nodes, Salpeter
    1-100, 5 % binaries with their own orbits) for 1/256 time unit with
    kira's call mix: binary centres of mass ask for perturber lists at every
    step (gamma = 1e-7, pert_freq 1) and make a perturbation pair call; a
    seeded sample of force calls and lists is checked."""
    from starlab_amd import ics
    from starlab_amd.g6 import GRAPE6
    N = 262144
    g = GRAPE6()
    assert g.open() == 0, g.last_error()
    idx = np.arange(N, dtype=np.int32)
    z = np.zeros((N, 3))

    def dev_pot(m, x):
        g.set_j_particles(idx, idx, np.zeros(N), np.ones(N), m, z, z, z, z, x)
        assert g.self_sweep(N, 0, N, 0.0, 0.0) == 0
        return 0.5 * float((m * g.read_results(N)[2]).sum())

    c = ics.king_cluster(N, 42, potential_fn=dev_pot)
    g.close()
    bins = dict(node=c["bin_index"], m1=c["m1"], dx=c["dx"], dv=c["dv"])
    chk = CallMixChecker(0.0, npipes=2048, every=97, max_i=16, seed=3, list_every=211)
    t0 = time.perf_counter()
    k = KiraCallMix(c["m"], c["x"], c["v"]).run(2.0 ** -8, binaries=bins, observer=chk)
    print(f"config 3 call mix to 1/256: {k.stats} wall {time.perf_counter() - t0:.1f}s",
          flush=True)
    _check_bars(chk)
    s = k.stats
    assert s["perturb_calls"] > 0 and s["list_gets"] > 0 and chk.list_checked > 0
    assert chk.kinds[CALL_PERTURB] == s["perturb_calls"]
    assert np.all(k.t == 2.0 ** -8)
