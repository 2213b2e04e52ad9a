"""kira's GRAPE-6 call mix in native code (bs_run_kira, libkira_blockstep.so)
on the CPU, against a GRAPE-6 ABI library whose arithmetic is the FP64 oracle
(tests/oracle_g6abi.c, built here): every force call of every kind -- blocks,
single-node redos after list rounds, perturbation pairs of binaries, energy
sweeps, density rounds -- and every neighbour list is checked against the
oracle on the j-memory exactly as the driver sent it; the counters show that
each of kira's paths ran (list overflows of both kinds, redos, releases and
re-attaches, resets after injected errors).  The rules are those of
hdyn_grape6.C; the numbers each path must produce are the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from starlab_amd.blockstep import (CALL_BLOCK, CALL_DENSITY, CALL_ENERGY, CALL_LIST,
                                   CALL_PERTURB, CALL_REDO, KiraCallMix, NativeBlockHermite)
from tests.kira_backends import CallMixChecker

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("abi") / "liboracle_g6abi.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
                    f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'oracle')}",
                    os.path.join(ROOT, "tests", "oracle_g6abi.c"),
                    os.path.join(ROOT, "oracle", "g6_oracle.c"), "-lm", "-o", str(out)],
                   check=True)
    return str(out)


def binaries(m, seed, count, a=1e-3):
    """count binary centre-of-mass nodes: component 1 holds 40 % of the node's
    mass, circular relative orbit of separation a, random orientation."""
    rng = np.random.default_rng(seed)
    node = np.sort(rng.choice(len(m), count, replace=False))
    dx = rng.normal(size=(count, 3))
    dx *= (a / np.linalg.norm(dx, axis=1))[:, None]
    ax = np.cross(dx, rng.normal(size=(count, 3)))
    ax /= np.linalg.norm(ax, axis=1)[:, None]
    dv = ax * np.sqrt(m[node] / a)[:, None]
    return dict(node=node, m1=0.4 * m[node], dx=dx, dv=dv)


@pytest.mark.parametrize("npipes", [2048, 48])
def test_every_call_matches_the_oracle(shim, npipes, monkeypatch):
    monkeypatch.setenv("ORACLE_G6ABI_NPIPES", str(npipes))
    m, x, v = oracle.makeplummer(256, 42)
    chk = CallMixChecker(0.0, npipes=npipes)
    k = KiraCallMix(m, x, v, eta=0.1).run(0.125, grape_lib=shim, star_radius=0.02,
                                          binaries=binaries(m, 1, 8), dt_log=1.0 / 32,
                                          densities=True, release_every=40, observer=chk)
    s = k.stats
    print(s, dict(chk.kinds), dict(chk.read_status))
    # the oracle library computes with the oracle: equality is exact
    assert chk.checked_blocks == s["force_calls"] and chk.nn_mismatch == 0
    assert chk.worst["acc_norm"] == 0 and chk.worst["jerk_norm"] == 0
    assert chk.worst["pot_rel"] == 0
    assert chk.list_checked == s["list_gets"] > 1000 and chk.list_mismatch == []
    for kind in (CALL_BLOCK, CALL_REDO, CALL_PERTURB, CALL_ENERGY, CALL_DENSITY, CALL_LIST):
        assert chk.kinds[kind] > 0, kind
    assert s["perturb_calls"] == chk.kinds[CALL_PERTURB]
    assert s["energy_sweeps"] == s["density_rounds"] == 5 and len(k.energies) == 5
    assert s["reopens"] >= 7 and s["full_reloads"] >= 10 and s["nn_from_lists"] > 0
    assert np.all(np.abs(k.energies / k.energies[0] - 1) < 1e-5)
    assert (k.density > 0).sum() > 200
    assert np.all(k.t == 0.125)


def test_without_extras_the_integration_is_bs_run(shim):
    """With no radii, binaries, logs or releases bs_run_kira makes kira's
    plain per-block sequence: bitwise the trajectory of bs_run."""
    m, x, v = oracle.makeplummer(200, 11)
    a = NativeBlockHermite(m, x, v, eta=0.1, eps2=1e-4).run(0.25, grape_lib=shim)
    b = KiraCallMix(m, x, v, eta=0.1, eps2=1e-4).run(0.25, grape_lib=shim)
    assert a.blocks == b.blocks and a.particle_steps == b.particle_steps
    for name in ("x", "v", "a", "j", "pot", "dt"):
        np.testing.assert_array_equal(getattr(b, name), getattr(a, name), err_msg=name)
    assert b.stats["list_reads"] == b.stats["perturb_calls"] == b.stats["resets"] == 0


def test_recovery_from_injected_errors(shim, monkeypatch):
    """lasthalf2 errors, list-read errors and an nn < 0 result (kira's error
    42) each make the driver reset, reload every j and retry
    (hdyn_grape6.C:405-438, 853-900, 2110-2166): the run ends bitwise equal to
    a clean one."""
    m, x, v = oracle.makeplummer(128, 7)
    kw = dict(grape_lib=shim, star_radius=0.02, dt_log=1.0 / 32)
    clean = KiraCallMix(m, x, v).run(0.0625, **kw)
    monkeypatch.setenv("ORACLE_G6ABI_FAIL_LASTHALF", "3")
    monkeypatch.setenv("ORACLE_G6ABI_FAIL_READNL", "2")
    monkeypatch.setenv("ORACLE_G6ABI_FAIL_NN", "1")
    chk = CallMixChecker(0.0, npipes=64)
    bad = KiraCallMix(m, x, v).run(0.0625, observer=chk, **kw)
    assert clean.stats["resets"] == 0
    assert bad.stats["resets"] == 6 and bad.stats["error42"] == 1
    assert chk.errors == 4 and chk.read_status[-1] == 2
    for name in ("x", "v", "a", "j", "pot", "dt"):
        np.testing.assert_array_equal(getattr(bad, name), getattr(clean, name), err_msg=name)
    np.testing.assert_array_equal(bad.energies, clean.energies)


def test_persistent_error_ends_the_run(shim, monkeypatch):
    """Six failures in a row exhaust kira's five resets (NRETRY): the run
    fails with an error instead of continuing on bad forces."""
    m, x, v = oracle.makeplummer(64, 7)
    monkeypatch.setenv("ORACLE_G6ABI_FAIL_LASTHALF", "6")
    with pytest.raises(RuntimeError, match="persistent error"):
        KiraCallMix(m, x, v).run(0.0625, grape_lib=shim)


def test_list_store_overflow_halves_radii_and_redoes(shim, monkeypatch):
    """Lists longer than the library's store make g6_read_neighbour_list_
    return > 0: kira moves the nodes without lists first, halves the radii
    and redoes the rest (sort_nodes_and_reduce_rnb, hdyn_grape6.C:2192-2255)."""
    monkeypatch.setenv("ORACLE_G6ABI_NB_CAP", "24")
    m, x, v = oracle.makeplummer(512, 9)
    chk = CallMixChecker(0.0, npipes=64)
    k = KiraCallMix(m, x, v).run(2.0 ** -6, grape_lib=shim, star_radius=0.3, observer=chk)
    s = k.stats
    print(s)
    assert s["read_overflows"] > 0 and chk.read_status[1] == s["read_overflows"]
    assert chk.list_checked == s["list_gets"] > 0 and chk.list_mismatch == []
    assert chk.worst["acc_norm"] == 0 and chk.nn_mismatch == 0


def test_long_lists_shrink_and_redo_single_nodes(shim, monkeypatch):
    """Lists of 4096 or more entries make g6_get_neighbour_list_ report
    overflow: kira shrinks that radius x 0.25, redoes the node alone (x 0.8
    per further overflow) and then the rest of its chunk
    (get_neighbors_and_adjust_h2, get_coll_and_perturbers,
    hdyn_grape6.C:2277-2998)."""
    monkeypatch.setenv("ORACLE_G6ABI_NPIPES", "48")
    m, x, v = oracle.makeplummer(4200, 3)
    chk = CallMixChecker(0.0, npipes=48, every=200, max_i=4, list_every=53)
    k = KiraCallMix(m, x, v).run(2.0 ** -10, grape_lib=shim, star_radius=3.0, observer=chk)
    s = k.stats
    print(s)
    assert s["get_overflows"] > 1000 and s["redo_calls"] >= s["get_overflows"]
    assert chk.list_checked > 100 and chk.list_mismatch == []
    assert chk.worst["acc_norm"] == 0 and chk.nn_mismatch == 0


def test_bad_options_are_errors(shim):
    m, x, v = oracle.makeplummer(16, 5)
    with pytest.raises(RuntimeError, match="bad binary"):
        KiraCallMix(m, x, v).run(0.125, grape_lib=shim,
                                 binaries=dict(node=[3], m1=[m[3] * 2], dx=[[1e-3, 0, 0]],
                                               dv=[[0, 1, 0]]))
    with pytest.raises(RuntimeError, match="bad options"):
        KiraCallMix(m, x, v).run(0.125, grape_lib=shim, star_radius=-1.0)
