"""The oracle against committed golden vectors of the reference's own IEEE
GRAPE-4 simulator (tests/golden/harp3_plummer300.npz, written by
tests/golden/make_harp3_golden.py from harp3fastsim.c compiled out of the
Starlab tree).  These pin the j predictor (4th order, GRAPE form) and the pair
law (acc, jerk) plus neighbour-list membership bit for bit, with or without
the reference tree present."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "harp3_plummer300.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD) as f:
        return {k: f[k] for k in f.files}


def test_predictor_matches_reference_simulator(gold):
    g = gold
    n = len(g["m"])
    xp, vp = oracle.predict(g["x"], g["v"], g["a2"], g["j6"], np.zeros((n, 3)), g["tj"],
                            float(g["ti"]), 4)
    np.testing.assert_array_equal(xp, g["xp"])
    np.testing.assert_array_equal(vp, g["vp"])


@pytest.mark.parametrize("e", [0, 1])
def test_force_and_lists_match_reference_simulator(gold, e):
    g = gold
    eps2 = float(g["eps2"][e])
    for s, sel in enumerate(g["sels"]):
        a, j, _, c, lists = oracle.g4_force(g["xp"][sel], g["vp"][sel], eps2, g["h2"], g["xp"],
                                            g["vp"], g["m"])
        np.testing.assert_array_equal(a, g[f"acc_{e}_{s}"])
        np.testing.assert_array_equal(j, g[f"jerk_{e}_{s}"])
        for k in range(3):
            np.testing.assert_array_equal(lists[k, :c[k]], g[f"list_{e}_{s}_{k}"])
    # the fixture is not trivial: some lists are non-empty, h2 = 0 gives none
    assert any(len(g[f"list_{e}_{s}_0"]) > 1 for s in range(len(g["sels"])))
    assert all(len(g[f"list_{e}_{s}_2"]) == 0 for s in range(len(g["sels"])))
