"""GPU parity of the energy and density fast paths of libg6.so
(include/g6x.h: g6x_potential, g6x_densities) through the C ABI.

Bars: pot relative error <= 1e-6 against the FP64 oracle (the force path's
bar) and total potential energy 1/2 sum m pot within 1e-6 relative of the
oracle's pairwise sum; k-th neighbour distances bit-identical to the oracle
(FP64, same operation order) and densities within 1e-13 relative (the
device's pow against the host's).
"""
import numpy as np
import pytest

import oracle
from tests.gpu_common import POT_TOL, load_j, plummer_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from starlab_amd.g6 import GRAPE6
    dev = GRAPE6()
    assert dev.open() == 0, dev.last_error()
    yield dev
    dev.close()


def _predict(s, ti):
    return oracle.predict(s["x"], s["v"], s["a2"], s["j6"], s["k18"], s["tj"], ti, 4)


@pytest.mark.parametrize("eps2", [0.0, 1e-6])
def test_potential_matches_oracle_and_energy(g, eps2):
    n = 8192
    s = plummer_state(n, seed=31)
    load_j(g, s)
    ti = 2.0 ** -11
    xp, vp = _predict(s, ti)
    pot = g.potential(n, 0, n, ti, eps2)
    _, _, rpot, _, _ = oracle.force(xp, vp, s["idx"], xp, vp, s["m"], s["idx"], eps2)
    rel = np.abs(pot - rpot) / np.abs(rpot)
    print("potential", eps2, rel.max(), flush=True)
    assert rel.max() <= POT_TOL
    ep, _ = oracle.energies(s["m"], xp, vp, eps2)
    assert abs(0.5 * (s["m"] * pot).sum() - ep) <= POT_TOL * abs(ep)
    # the full force path gives the same pot to the same bar
    assert g.self_sweep(n, 0, n, ti, eps2) == 0
    _, _, pfull, _ = g.read_results(n)
    assert (np.abs(pot - pfull) / np.abs(rpot)).max() <= 2 * POT_TOL


def test_potential_subrange_and_energy_snapshot(g):
    """kira's energy snapshot: every t_j = t_i = 0, so prediction is the
    identity; an i-subrange gives the same pot as the full sweep."""
    n = 3000
    s = plummer_state(n, seed=5)
    s = dict(s, tj=np.zeros(n))
    load_j(g, s)
    pot = g.potential(n, 0, n, 0.0, 0.0)
    sub = g.potential(n, 1000, 777, 0.0, 0.0)
    _, _, rpot, _, _ = oracle.force(s["x"], s["v"], s["idx"], s["x"], s["v"], s["m"],
                                    s["idx"], 0.0)
    assert (np.abs(pot - rpot) / np.abs(rpot)).max() <= POT_TOL
    assert (np.abs(sub - rpot[1000:1777]) / np.abs(rpot[1000:1777])).max() <= POT_TOL


@pytest.mark.parametrize("k", [2, 12, 17, 32])
def test_densities_match_oracle(g, k):
    n = 6000
    s = plummer_state(n, seed=44)
    rng = np.random.default_rng(k)
    s = dict(s, m=rng.uniform(0.5, 1.5, n) / n)           # unequal masses
    load_j(g, s)
    ti = 2.0 ** -12
    xp, _ = _predict(s, ti)
    rho, dk2 = g.densities(n, 0, n, ti, k)
    rrho, rdk2 = oracle.density(s["m"], xp, k)
    np.testing.assert_array_equal(dk2, rdk2)
    rel = np.abs(rho - rrho) / rrho
    print("densities k", k, rel.max(), flush=True)
    assert rel.max() <= 1e-13


def test_densities_lattice_ties_and_subrange(g):
    """A cubic lattice (many equal distances): with equal masses the tied
    shells give the closed form; a subrange equals the full sweep."""
    a = 0.05
    gr = np.arange(-6, 7) * a
    x = np.array(np.meshgrid(gr, gr, gr, indexing="ij")).reshape(3, -1).T.copy()
    n = len(x)
    z = np.zeros((n, 3))
    idx = np.arange(n, dtype=np.int32)
    m = np.full(n, 1.0 / n)
    assert g.set_j_particles(idx, idx, np.zeros(n), np.ones(n), m, z, z, z, z, x) == 0
    rho, dk2 = g.densities(n, 0, n, 0.0, 12)
    rrho, rdk2 = oracle.density(m, x, 12)
    np.testing.assert_array_equal(dk2, rdk2)
    assert (np.abs(rho - rrho) / rrho).max() <= 1e-13
    c = n // 2
    want = 11 * m[0] / ((4.0 / 3.0) * np.pi * (2 * a * a) ** 1.5)
    assert abs(rho[c] - want) <= 1e-12 * want
    sub, subk = g.densities(n, 100, 500, 0.0, 12)
    np.testing.assert_array_equal(subk, dk2[100:600])
    np.testing.assert_array_equal(sub, rho[100:600])


def test_densities_full_size_sample(g):
    """N = 65536 (the force-only workload's size), every particle on the
    device; the oracle checks a 512-particle sample."""
    n = 65536
    s = plummer_state(n, seed=65536)
    load_j(g, s)
    xp, _ = _predict(s, 0.0)
    rho, dk2 = g.densities(n, 0, n, 0.0, 12)
    sel = 4096
    rrho, rdk2 = oracle.density(s["m"], xp, 12, sel, 512)
    np.testing.assert_array_equal(dk2[sel:sel + 512], rdk2)
    assert (np.abs(rho[sel:sel + 512] - rrho) / rrho).max() <= 1e-13
    assert np.all(rho > 0)


def test_bad_arguments_fail_cleanly(g):
    with pytest.raises(RuntimeError):
        g.densities(10, 0, 10, 0.0, 10)      # k >= nj
    with pytest.raises(RuntimeError):
        g.densities(100, 0, 100, 0.0, 40)    # k > 32
    with pytest.raises(RuntimeError):
        g.potential(100, 90, 20, 0.0, 0.0)   # range past nj
