"""The native block-step driver (libkira_blockstep.so) on the CPU: driven
against a GRAPE-6 ABI library whose arithmetic is the FP64 oracle
(tests/oracle_g6abi.c, built here), it must reproduce the Python driver
(starlab_amd/blockstep.py) on the same arithmetic bit for bit -- same
blocks, same steps, same final state."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from starlab_amd.blockstep import BlockHermite, NativeBlockHermite
from tests.kira_backends import OracleAbiBackend

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


@pytest.mark.parametrize("npipes", [2048, 48])
def test_native_equals_python_driver_bitwise(shim, npipes, monkeypatch):
    monkeypatch.setenv("ORACLE_G6ABI_NPIPES", str(npipes))
    m, x, v = oracle.makeplummer(200, 11)
    py = BlockHermite(m, x, v, eta=0.1, eps2=1e-4)
    py.run(OracleAbiBackend(), 0.25)
    nat = NativeBlockHermite(m, x, v, eta=0.1, eps2=1e-4).run(0.25, grape_lib=shim)
    assert nat.blocks == py.blocks and nat.particle_steps == py.particle_steps
    for name in ("x", "v", "a", "j", "pot", "dt"):
        np.testing.assert_array_equal(getattr(nat, name), getattr(py, name), err_msg=name)
    assert np.all(nat.t == 0.25)
    assert nat.abi_seconds > 0 and nat.seconds >= nat.abi_seconds
    # bs_timers: every phase ran, the parts fit in the whole, the ABI ones in
    # the ABI time
    ph = nat.phase_seconds
    assert all(ph[k] > 0 for k in ("send_j", "firsthalf", "lasthalf", "predict", "correct"))
    parts = sum(ph[k] for k in ("send_j", "firsthalf", "lasthalf", "predict", "correct"))
    assert parts <= ph["total"] and abs(ph["total"] - nat.seconds) < 0.05 * nat.seconds + 1e-3
    assert ph["firsthalf"] + ph["lasthalf"] <= nat.abi_seconds


def test_observer_sees_blocks_and_jmemory(shim):
    """Blocks arrive in time order with ascending indices; j-memory holds
    every particle's last sent state, which the oracle turns back into the
    block's forces exactly."""
    m, x, v = oracle.makeplummer(96, 3)
    seen = []

    def obs(jm, t, idx, xi, vi, acc, jerk, pot, nn):
        assert np.all(np.diff(idx) > 0) and np.all(jm["t"] <= t)
        if len(seen) % 7 == 0:
            xp, vp = oracle.predict(jm["x"], jm["v"], jm["acc"] / 2.0, jm["jerk"] / 6.0,
                                    jm["k18"], jm["t"], t, 4)
            ra, rj, rp, _, _ = oracle.force(xi, vi, idx, xp, vp, jm["m"],
                                            np.arange(len(jm["m"]), dtype=np.int32), 1e-4)
            np.testing.assert_array_equal(acc, ra)
            np.testing.assert_array_equal(jerk, rj)
            np.testing.assert_array_equal(pot, rp)
        seen.append(t)
        return False

    nat = NativeBlockHermite(m, x, v, eta=0.1, eps2=1e-4).run(0.125, grape_lib=shim,
                                                             observer=obs)
    assert len(seen) == nat.blocks > 10 and seen == sorted(seen)


def test_observer_can_stop_and_errors_surface(shim, tmp_path):
    m, x, v = oracle.makeplummer(32, 5)
    with pytest.raises(RuntimeError, match="stopped by the observer"):
        NativeBlockHermite(m, x, v).run(0.125, grape_lib=shim, observer=lambda *a: True)
    with pytest.raises(ZeroDivisionError):
        NativeBlockHermite(m, x, v).run(0.125, grape_lib=shim, observer=lambda *a: 1 / 0)
    with pytest.raises(RuntimeError, match="dlopen failed"):
        NativeBlockHermite(m, x, v).run(0.125, grape_lib=str(tmp_path / "missing.so"))


@pytest.mark.parametrize("kw,t_end", [(dict(eta=0.0), 0.25), (dict(eps2=-1.0), 0.25),
                                      (dict(dt_max=0.0), 0.25), ({}, 0.0),
                                      (dict(eta=float("nan")), 0.25), ({}, float("inf"))])
def test_bad_scalars_are_errors(shim, kw, t_end):
    """Arguments that would give zero or non-finite steps (a run that never
    ends) are rejected before the library is opened."""
    m, x, v = oracle.makeplummer(16, 5)
    with pytest.raises(RuntimeError, match="t_end > 0"):
        NativeBlockHermite(m, x, v, **{"eta": 0.1, "eps2": 1e-4, **kw}).run(t_end, grape_lib=shim)
