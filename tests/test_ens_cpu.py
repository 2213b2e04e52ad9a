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
