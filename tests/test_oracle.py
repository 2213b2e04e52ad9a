"""Pins of the CPU oracle (oracle/g6_oracle.c) -- no GPU.

Each check is chosen so that a plausible slip in the oracle (dropped term,
wrong sign, swapped operand, wrong index) fails at least one of them:
closed forms, conservation laws, finite differences, long-double brute force,
and the reference's own IEEE GRAPE simulator compiled from its source
(oracle/_ref, src/harpsims/harp3fastsim.c).
"""
import math

import numpy as np
import pytest

import oracle


# --------------------------------------------------------------------------
# ran2 (src/std/randinter.C:257-324)
# --------------------------------------------------------------------------

def _ran2_bigint(seed, n):
    """ran2 with exact big-integer modular products instead of Schrage's
    decomposition: pins the overflow-free arithmetic of the C version."""
    M1, M2, A1, A2 = 2147483563, 2147483399, 40014, 40692
    NTAB = 32
    NDIV = 1 + (M1 - 1) // NTAB
    s = max(abs(seed), 1)
    s2 = s
    v = [0] * NTAB
    for t in range(NTAB + 7, -1, -1):
        s = (A1 * s) % M1
        if t < NTAB:
            v[t] = s
    y = v[0]
    out = []
    for _ in range(n):
        s = (A1 * s) % M1
        s2 = (A2 * s2) % M2
        k = y // NDIV
        y = v[k] - s2
        v[k] = s
        if y < 1:
            y += M1 - 1
        u = (1.0 / M1) * y
        out.append(min(u, 1.0 - 1e-14))
    return np.array(out)


def test_ran2_matches_bigint_restatement():
    for seed in (1, 42, 12345, 2147483000):
        np.testing.assert_array_equal(oracle.ran2_sequence(seed, 500), _ran2_bigint(seed, 500))


def test_ran2_statistics_and_determinism():
    u = oracle.ran2_sequence(7, 100000)
    assert 0.0 < u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005
    hist, _ = np.histogram(u, bins=20, range=(0, 1))
    chi2 = ((hist - 5000.0) ** 2 / 5000.0).sum()
    assert chi2 < 45          # 19 dof, p ~ 1e-3
    np.testing.assert_array_equal(u[:10], oracle.ran2_sequence(7, 10))
    assert not np.array_equal(u[:10], oracle.ran2_sequence(8, 10))


def test_package_ran2_equals_oracle():
    from starlab_amd.ics import Ran2
    r = Ran2(42)
    got = np.array([r.randinter(0.0, 1.0) for _ in range(2000)])
    np.testing.assert_array_equal(got, oracle.ran2_sequence(42, 2000))


# --------------------------------------------------------------------------
# Plummer initial conditions (makeplummer.C:55-214)
# --------------------------------------------------------------------------

def test_plummer_invariants():
    m, x, v = oracle.makeplummer(1024, 42)
    assert abs(m.sum() - 1.0) < 1e-14
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-14
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-14
    ep, ek = oracle.energies(m, x, v)
    assert abs(ep + ek + 0.25) < 1e-13
    assert abs(ek / -ep - 0.5) < 1e-13
    # half-mass radius of a Plummer sphere in virial units: 0.769 (+- sampling)
    r = np.sort(np.sqrt((x * x).sum(1)))
    assert 0.65 < r[len(r) // 2] < 0.9


def test_package_plummer_matches_oracle():
    from starlab_amd.ics import makeplummer
    m0, x0, v0 = oracle.makeplummer(256, 42, scale=False)
    m1, x1, v1 = makeplummer(256, 42, scale=False)
    np.testing.assert_array_equal(m1, m0)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(v1, v0)
    m0, x0, v0 = oracle.makeplummer(256, 42)
    m1, x1, v1 = makeplummer(256, 42)
    np.testing.assert_allclose(x1, x0, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(v1, v0, rtol=1e-12, atol=1e-13)


# --------------------------------------------------------------------------
# predictor
# --------------------------------------------------------------------------

def _random_j(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    a = rng.normal(size=(n, 3))
    j = rng.normal(size=(n, 3))
    s = rng.normal(size=(n, 3))
    tj = rng.uniform(0, 2.0 ** -10, n)
    return x, v, a, j, s, tj


def test_predictor_identity_at_dt0():
    x, v, a, j, s, tj = _random_j(100)
    xp, vp = oracle.predict(x, v, a / 2, j / 6, s / 18, np.zeros(100), 0.0, 5)
    np.testing.assert_array_equal(xp, x)
    np.testing.assert_array_equal(vp, v)


def test_predictor_taylor_polynomials():
    # 4th order reproduces a cubic trajectory, 5th order a quartic (exact
    # polynomials, rounding only)
    x, v, a, j, s, tj = _random_j(200)
    ti = 0.05
    dt = (ti - tj)[:, None]
    cubic_x = x + v * dt + a * dt ** 2 / 2 + j * dt ** 3 / 6
    cubic_v = v + a * dt + j * dt ** 2 / 2
    xp, vp = oracle.predict(x, v, a / 2, j / 6, s / 18, tj, ti, 4)
    np.testing.assert_allclose(xp, cubic_x, rtol=0, atol=1e-15)
    np.testing.assert_allclose(vp, cubic_v, rtol=0, atol=1e-15)
    quart_x = cubic_x + s * dt ** 4 / 24
    quart_v = cubic_v + s * dt ** 3 / 6
    xp, vp = oracle.predict(x, v, a / 2, j / 6, s / 18, tj, ti, 5)
    np.testing.assert_allclose(xp, quart_x, rtol=0, atol=1e-15)
    np.testing.assert_allclose(vp, quart_v, rtol=0, atol=1e-15)
    # and the k18 term is really used
    assert np.abs(xp - cubic_x).max() > 1e-10


@pytest.mark.skipif(not oracle.Harp3Sim.available(), reason="reference tree absent at build time")
def test_predictor_bitwise_vs_reference_simulator():
    h = oracle.Harp3Sim()
    x, v, a, j, s, tj = _random_j(300)
    m = np.full(300, 1.0 / 300)
    h.set_j(x, v, a / 2, j / 6, m, tj)
    ti = 0.0009
    h.calc(ti, 300, x[:1], v[:1], 0.0, np.zeros(1))
    hx, hv = h.predicted(300)
    xp, vp = oracle.predict(x, v, a / 2, j / 6, s / 18, tj, ti, 4)
    np.testing.assert_array_equal(xp, hx)
    np.testing.assert_array_equal(vp, hv)


# --------------------------------------------------------------------------
# force law (hdyn_ev.C:2019-2036)
# --------------------------------------------------------------------------

def _force_all(m, x, v, eps2):
    n = len(m)
    idx = np.arange(n, dtype=np.int32)
    return oracle.force(x, v, idx, x, v, m, idx, eps2)


def test_two_body_closed_form():
    m = np.array([0.3, 0.7])
    x = np.array([[0.1, -0.2, 0.3], [0.5, 0.4, -0.1]])
    v = np.array([[0.0, 0.1, -0.2], [0.3, -0.1, 0.2]])
    eps2 = 0.01
    acc, jerk, pot, nn, dnn2 = _force_all(m, x, v, eps2)
    d = x[1] - x[0]
    w = v[1] - v[0]
    r2 = d @ d + eps2
    r = math.sqrt(r2)
    np.testing.assert_allclose(acc[0], m[1] * d / r ** 3, rtol=1e-14)
    np.testing.assert_allclose(acc[1], -m[0] * d / r ** 3, rtol=1e-14)
    np.testing.assert_allclose(jerk[0], m[1] * (w / r ** 3 - 3 * (d @ w) * d / r ** 5), rtol=1e-13)
    np.testing.assert_allclose(pot, [-m[1] / r, -m[0] / r], rtol=1e-14)
    assert list(nn) == [1, 0]
    np.testing.assert_allclose(dnn2, [d @ d, d @ d], rtol=1e-15)


def test_newton_third_law_and_potential_energy():
    rng = np.random.default_rng(3)
    n = 300
    m = rng.uniform(0.1, 2.0, n)
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    acc, jerk, pot, _, _ = _force_all(m, x, v, 1e-4)
    scale = (m[:, None] * np.abs(acc)).sum(0)
    assert np.all(np.abs((m[:, None] * acc).sum(0)) < 1e-13 * scale)
    jscale = (m[:, None] * np.abs(jerk)).sum(0)
    assert np.all(np.abs((m[:, None] * jerk).sum(0)) < 1e-13 * jscale)
    ep, _ = oracle.energies(m, x, v, 1e-4)
    assert abs(0.5 * (m * pot).sum() - ep) < 1e-12 * abs(ep)


def test_translation_and_galilean_invariance():
    rng = np.random.default_rng(4)
    n = 128
    m = rng.uniform(0.5, 1.5, n)
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    a0, j0, p0, nn0, _ = _force_all(m, x, v, 1e-3)
    a1, j1, p1, nn1, _ = _force_all(m, x + np.array([0.25, -0.5, 0.125]),
                                    v + np.array([1.5, -2.0, 0.5]), 1e-3)
    np.testing.assert_allclose(a1, a0, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(j1, j0, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(p1, p0, rtol=1e-12)
    np.testing.assert_array_equal(nn1, nn0)


def test_jerk_is_time_derivative_of_acc():
    rng = np.random.default_rng(5)
    n = 64
    m = rng.uniform(0.5, 1.5, n)
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    h = 1e-5
    _, jerk, _, _, _ = _force_all(m, x, v, 1e-2)
    ap, _, _, _, _ = _force_all(m, x + h * v, v, 1e-2)
    am, _, _, _, _ = _force_all(m, x - h * v, v, 1e-2)
    fd = (ap - am) / (2 * h)
    np.testing.assert_allclose(jerk, fd, rtol=1e-6, atol=1e-6 * np.abs(jerk).max())


def test_long_double_brute_force():
    rng = np.random.default_rng(6)
    n = 200
    m = rng.uniform(0.5, 1.5, n) / n
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    idx = np.arange(n, dtype=np.int32)
    acc, jerk, pot, _, _ = _force_all(m, x, v, 1e-6)
    for i in (0, 17, 99, 199):
        a, j, p = oracle.force_ld(x[i], v[i], i, x, v, m, idx, 1e-6)
        np.testing.assert_allclose(acc[i], a, rtol=1e-12, atol=1e-14 * np.abs(a).max())
        np.testing.assert_allclose(jerk[i], j, rtol=1e-12, atol=1e-14 * np.abs(j).max())
        assert abs(pot[i] - p) < 1e-13 * abs(p)


def test_exclusion_is_by_index_not_position():
    # the perturbation path passes both binary components with the parent's
    # index (hdyn_grape6.C:3462-3468): that j is skipped, all others count
    rng = np.random.default_rng(7)
    n = 50
    m = rng.uniform(0.5, 1.5, n)
    xj = rng.normal(size=(n, 3))
    vj = rng.normal(size=(n, 3))
    jidx = np.arange(n, dtype=np.int32)
    xi = xj[[10, 10]] + np.array([[1e-3, 0, 0], [-1e-3, 0, 0]])
    vi = vj[[10, 10]]
    acc, _, pot, nn, _ = oracle.force(xi, vi, np.array([10, 10], np.int32), xj, vj, m, jidx, 0.0)
    keep = jidx != 10
    for k in range(2):
        d = xj[keep] - xi[k]
        r = np.sqrt((d * d).sum(1))
        np.testing.assert_allclose(acc[k], (m[keep, None] * d / r[:, None] ** 3).sum(0), rtol=1e-12)
        np.testing.assert_allclose(pot[k], -(m[keep] / r).sum(), rtol=1e-13)
        assert nn[k] == jidx[keep][np.argmin(r)]


def test_nn_ties_pick_lowest_index():
    # a cubic lattice: every interior point has 6 equidistant neighbours
    g = np.arange(4, dtype=float)
    x = np.array([[a, b, c] for a in g for b in g for c in g])
    n = len(x)
    m = np.full(n, 1.0 / n)
    v = np.zeros_like(x)
    _, _, _, nn, dnn2 = _force_all(m, x, v, 0.0)
    for i in range(n):
        d2 = ((x - x[i]) ** 2).sum(1)
        d2[i] = np.inf
        assert nn[i] == int(np.flatnonzero(d2 == d2.min())[0])
        assert dnn2[i] == 1.0


@pytest.mark.skipif(not oracle.Harp3Sim.available(), reason="reference tree absent at build time")
def test_force_vs_reference_simulator():
    """acc and jerk against the reference's IEEE GRAPE-4 simulator
    (harp3fastsim.c:268-322); it softens with 2*eps2_in, so pass eps2/2."""
    h = oracle.Harp3Sim()
    rng = np.random.default_rng(8)
    n = 900
    m = rng.uniform(0.5, 1.5, n) / n
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    zero = np.zeros((n, 3))
    h.set_j(x, v, zero, zero, m, np.zeros(n))
    idx = np.arange(n, dtype=np.int32)
    eps2 = 1e-3
    for s in range(0, 9, 3):
        sl = slice(s, s + 3)
        ha, hj, _ = h.calc(0.0, n, x[sl], v[sl], eps2, np.zeros(3))
        acc, jerk, _, _, _ = oracle.force(x[sl], v[sl], idx[sl], x, v, m, idx, eps2)
        np.testing.assert_allclose(acc, ha, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(jerk, hj, rtol=1e-11, atol=1e-13)


# --------------------------------------------------------------------------
# neighbour lists
# --------------------------------------------------------------------------

def test_neighbours_brute_force():
    rng = np.random.default_rng(9)
    n = 400
    x = rng.uniform(-1, 1, size=(n, 3))
    idx = np.arange(n, dtype=np.int32)
    h2 = rng.uniform(0, 0.1, n)
    h2[::7] = 0.0
    for incl in (0, 1):
        counts, lists = oracle.neighbours(x, h2, idx, incl, x, idx, 64)
        for i in range(n):
            d2 = ((x - x[i]) ** 2).sum(1)
            want = [j for j in range(n) if d2[j] < h2[i] and (incl or j != i)] if h2[i] > 0 else []
            assert counts[i] == len(want)
            assert list(lists[i, :min(64, len(want))]) == want[:64]


@pytest.mark.skipif(not oracle.Harp3Sim.available(), reason="reference tree absent at build time")
def test_neighbours_vs_reference_simulator():
    """The simulator's list is {j : r^2 + 2 eps2_in < 2 h2_in} in ascending j
    with self kept (r^2 = 0 < h2): eps2 = 0 and h2 halved give {r^2 < h2}."""
    h = oracle.Harp3Sim()
    rng = np.random.default_rng(10)
    n = 700
    x = rng.uniform(-1, 1, size=(n, 3))
    v = rng.normal(size=(n, 3))
    m = np.full(n, 1.0 / n)
    zero = np.zeros((n, 3))
    h.set_j(x, v, zero, zero, m, np.zeros(n))
    idx = np.arange(n, dtype=np.int32)
    h2 = np.array([0.05, 0.02, 0.1])
    sl = slice(100, 103)
    _, _, lists = h.calc(0.0, n, x[sl], v[sl], 0.0, h2)
    counts, ol = oracle.neighbours(x[sl], h2, idx[sl], 1, x, idx, 1024)
    for k in range(3):
        assert counts[k] == len(lists[k])
        np.testing.assert_array_equal(ol[k, :counts[k]], lists[k])
