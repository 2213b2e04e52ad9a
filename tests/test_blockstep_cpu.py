"""The block-step driver on the CPU (FP64 host backend) -- no GPU."""
import numpy as np

import oracle
from starlab_amd.blockstep import BlockHermite, first_step
from tests.kira_backends import HostBackend


def test_first_step_is_power_of_two_and_bounded():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(50, 3)); j = rng.normal(size=(50, 3)); pot = -rng.uniform(1, 2, 50)
    dt = first_step(0.1, a, j, pot, 0.125)
    assert np.all(dt <= 0.1 / 16)
    assert np.all(np.log2(dt) == np.round(np.log2(dt)))


def test_small_plummer_energy_and_schedule():
    m, x, v = oracle.makeplummer(64, 42)
    b = BlockHermite(m, x, v, eta=0.05, eps2=1e-4)
    e0 = b.energy(lambda m, x, v, e: oracle.energies(m, x, v, e))
    seen = []
    b.run(HostBackend(), 0.25, on_block=lambda s, t, blk, *r: seen.append((t, len(blk))))
    assert np.all(b.t == 0.25)                 # every particle lands on t_end
    ts = [t for t, _ in seen]
    assert ts == sorted(ts) and len(set(ts)) == len(ts)
    # block times are multiples of the members' steps (power-of-two schedule)
    assert all(abs(t * 2 ** 40 - round(t * 2 ** 40)) == 0 for t in ts)
    e1 = b.energy(lambda m, x, v, e: oracle.energies(m, x, v, e))
    assert abs((e1 - e0) / e0) < 1e-6
    assert b.particle_steps < len(seen) * 64   # individual steps, not shared


def test_scheduler_picks_the_min_time_set():
    """Each block is exactly {i : t_i + dt_i = min_k (t_k + dt_k)}, ascending."""
    m, x, v = oracle.makeplummer(128, 7)
    b = BlockHermite(m, x, v, eta=0.05, eps2=1e-4)
    seen = []

    def check(s, t, blk, *r):
        tn = s.t + s.dt
        want = np.flatnonzero(tn == tn.min())
        assert tn.min() == t
        np.testing.assert_array_equal(blk, want)
        seen.append(len(blk))

    b.run(HostBackend(), 0.125, on_block=check)
    assert len(seen) > 10 and min(seen) < 128
