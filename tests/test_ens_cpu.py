"""Ensemble workloads and the shared-step Hermite oracle -- no GPU."""
import math

import numpy as np

import oracle
from starlab_amd.ens import binary_single, plummer_cluster


def test_plummer_cluster_invariants():
    rng = np.random.default_rng(3)
    for n in (32, 500, 2048):
        m, x, v = plummer_cluster(n, rng)
        assert abs(m.sum() - 1) < 1e-12
        assert np.abs((m[:, None] * x).sum(0)).max() < 1e-12
        assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
        ep, ek = oracle.energies(m, x, v)
        q = ek / -ep
        assert 0.35 < q < 0.65          # virial within sampling noise for small n
        if n == 2048:
            assert abs(ep + ek + 0.25) < 0.03


def test_binary_single_conventions():
    rng = np.random.default_rng(5)
    for _ in range(20):
        m, x, v = binary_single(rng)
        assert abs(m[:2].sum() - 1) < 1e-15
        assert np.abs((m[:, None] * x).sum(0)).max() < 1e-12
        assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
        # inner binary: semi-major axis 1
        d = x[1] - x[0]
        w = v[1] - v[0]
        eb = 0.5 * w @ w - 1.0 / np.linalg.norm(d)
        assert abs(-1.0 / (2 * eb) - 1.0) < 1e-10
        # projectile 20 length units away, moving inwards
        cm = (m[:2, None] * x[:2]).sum(0)
        rel = x[2] - cm
        relv = v[2] - (m[:2, None] * v[:2]).sum(0)
        assert abs(np.linalg.norm(rel) - 20.0) < 1e-10
        assert rel @ relv < 0


def _kepler_circular():
    m = np.array([0.5, 0.5])
    x = np.array([[-0.5, 0, 0], [0.5, 0, 0]], dtype=float)
    vc = math.sqrt(0.5 * 0.5 / 1.0)      # relative speed 1, each moves at 0.5
    v = np.array([[0, -0.5, 0], [0, 0.5, 0]], dtype=float)
    assert abs(vc - 0.5) < 1e-15
    return m, x, v


def test_oracle_integrator_kepler_period():
    m, x, v = _kepler_circular()
    period = 2 * math.pi            # a = 1, M = 1
    xf, vf, t, steps, e0, e1, rc = oracle.hermite_shared(m, x, v, period, 0.05, 0.0, dt_max=0.125)
    assert rc == 0 and abs(t - period) < 1e-12
    assert np.abs(xf - x).max() < 1e-6
    assert abs((e1 - e0) / e0) < 1e-8


def test_oracle_integrator_fourth_order():
    rng = np.random.default_rng(11)
    m, x, v = plummer_cluster(64, rng)
    errs = []
    for eta in (0.04, 0.02):
        _, _, t, steps, e0, e1, rc = oracle.hermite_shared(m, x, v, 0.25, eta, 1e-2)
        assert rc == 0 and t == 0.25
        errs.append(abs((e1 - e0) / e0))
    # halving eta shrinks steps ~2x: a 4th-order scheme gains ~2^4
    assert errs[1] < errs[0] / 5


def test_oracle_integrator_conserves_momentum():
    rng = np.random.default_rng(12)
    m, x, v = binary_single(rng)
    xf, vf, t, steps, e0, e1, rc = oracle.hermite_shared(m, x, v, 8.0, 0.05, 0.0)
    assert rc == 0
    assert np.abs((m[:, None] * vf).sum(0)).max() < 1e-12
    assert abs((e1 - e0) / e0) < 1e-5


# ---------------------------------------------------------------------------
# The block-step oracle (oracle.hermite_block), checker of the engine's block
# mode.  Pins: kira's block-step rules restated independently, the special
# cases where every particle keeps one step (then it is the shared-step
# oracle), orbit closure and fourth-order convergence.
# ---------------------------------------------------------------------------
def test_block_oracle_equals_kira_rules_restated():
    """starlab_amd/blockstep.py restates kira's block scheduling, corrector
    and step rules in vectorised numpy; on an FP64 host backend (the same
    predictor and pair law) it must give the oracle's trajectory bit for bit,
    with the same blocks and particle steps."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from kira_backends import HostBackend
    from starlab_amd.blockstep import BlockHermite
    rng = np.random.default_rng(11)
    for n in (16, 64, 200):
        m, x, v = plummer_cluster(n, rng)
        eps2 = (4.0 / n) ** 2
        xo, vo, t, nb, ps, e0, e1, rc = oracle.hermite_block(m, x, v, 0.25, 0.1, eps2)
        b = BlockHermite(m, x, v, eta=0.1, eps2=eps2)
        b.run(HostBackend(), 0.25)
        assert rc == 0 and t == 0.25
        assert (nb, ps) == (b.blocks, b.particle_steps)
        assert ps < n * nb                  # particles did take different steps
        np.testing.assert_array_equal(xo, b.x)
        np.testing.assert_array_equal(vo, b.v)


def test_block_oracle_is_the_shared_one_when_all_steps_agree():
    """Two equal masses on mirror-image orbits have bitwise equal step
    criteria, and a step cap below every particle's natural step forces one
    step on all: in both cases block steps are the shared step, bit for bit."""
    m, x, v = _kepler_circular()
    x[:, 0] *= 1.3                       # eccentric, still mirror-symmetric
    xs, vs, ts, steps, e0s, e1s, rcs = oracle.hermite_shared(m, x, v, 3.0, 0.05, 0.0)
    xb, vb, tb, nb, ps, e0b, e1b, rcb = oracle.hermite_block(m, x, v, 3.0, 0.05, 0.0)
    assert rcs == rcb == 0 and ts == tb == 3.0
    assert nb == steps and ps == 2 * steps
    np.testing.assert_array_equal(xb, xs)
    np.testing.assert_array_equal(vb, vs)
    assert e1b == e1s
    rng = np.random.default_rng(21)
    mm, xx, vv = plummer_cluster(32, rng)
    cap = 2.0 ** -12
    xs, vs, ts, steps, _, e1s, rcs = oracle.hermite_shared(mm, xx, vv, 1.0 / 16, 0.1, 1e-2, dt_max=cap)
    xb, vb, tb, nb, ps, _, e1b, rcb = oracle.hermite_block(mm, xx, vv, 1.0 / 16, 0.1, 1e-2, dt_max=cap)
    assert rcs == rcb == 0 and nb == steps == (1.0 / 16) / cap and ps == 32 * nb
    np.testing.assert_array_equal(xb, xs)
    assert e1b == e1s


def test_block_oracle_kepler_period_and_fourth_order():
    """An unequal-mass eccentric binary (each body on its own steps) closes
    after one period; on a Plummer cluster the energy error falls as eta^4."""
    m = np.array([0.8, 0.2])
    e_orb = 0.5
    rp = 1.0 - e_orb                                    # a = 1, M = 1, pericentre on x
    vp = math.sqrt((1 + e_orb) / rp)
    x = np.array([[-0.2 * rp, 0, 0], [0.8 * rp, 0, 0]])
    v = np.array([[0, -0.2 * vp, 0], [0, 0.8 * vp, 0]])
    period = 2 * math.pi
    xf, vf, t, nb, ps, e0, e1, rc = oracle.hermite_block(m, x, v, period, 0.02, 0.0)
    assert rc == 0 and abs(t - period) < 1e-12
    assert np.abs(xf - x).max() < 1e-5
    assert abs((e1 - e0) / e0) < 1e-8
    rng = np.random.default_rng(11)
    mm, xx, vv = plummer_cluster(64, rng)
    errs = []
    for eta in (0.04, 0.02):
        _, _, t, nb, ps, e0, e1, rc = oracle.hermite_block(mm, xx, vv, 0.25, eta, 1e-2)
        assert rc == 0 and t == 0.25
        errs.append(abs((e1 - e0) / e0))
    assert errs[1] < errs[0] / 5
