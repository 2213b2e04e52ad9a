"""Write tests/golden/harp3_plummer300.npz: input/output vectors of the
reference's own IEEE GRAPE-4 simulator (src/harpsims/harp3fastsim.c of the
Starlab tree, compiled from source into oracle/_ref by oracle/Makefile).

    python tests/golden/make_harp3_golden.py

Inputs: a 300-particle Plummer sphere (oracle.makeplummer, seed 7) with
seeded a/2, j/6 and j-times t_j spread over [0, 2^-6], predicted to
t_i = 2^-6.  Outputs: the simulator's predicted positions and velocities of
every j, and for three triples of i-particles (taken from j-memory) at
eps2 = 0 and 1e-4: acc, jerk and the neighbour lists (h2 = 0.02, 0.05, 0).
The fixture pins the oracle without the reference tree present
(tests/test_golden_cpu.py).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def main():
    sim = oracle.Harp3Sim()
    n = 300
    m, x, v = oracle.makeplummer(n, 7)
    rng = np.random.default_rng(7)
    a2 = rng.normal(scale=0.1, size=(n, 3))
    j6 = rng.normal(scale=0.01, size=(n, 3))
    tj = rng.uniform(0.0, 2.0 ** -6, n)
    ti = 2.0 ** -6
    sim.set_j(x, v, a2, j6, m, tj)
    sels = np.array([[5, 6, 7], [0, 150, 299], [42, 43, 250]], dtype=np.int32)
    h2 = np.array([0.02, 0.05, 0.0])
    out = {}
    for e, eps2 in enumerate((0.0, 1e-4)):
        for s, sel in enumerate(sels):
            # a calc at t_i predicts all j first; read the prediction back
            acc0, _, _ = sim.calc(ti, n, x[sel], v[sel], 2 * eps2, 2 * h2)
            xp, vp = sim.predicted(n)
            acc, jerk, lists = sim.calc(ti, n, xp[sel], vp[sel], 2 * eps2, 2 * h2)
            out[f"acc_{e}_{s}"] = acc
            out[f"jerk_{e}_{s}"] = jerk
            for k in range(3):
                out[f"list_{e}_{s}_{k}"] = np.asarray(lists[k], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "harp3_plummer300.npz"), m=m, x=x, v=v, a2=a2, j6=j6,
                        tj=tj, ti=np.array(ti), xp=xp, vp=vp, sels=sels, h2=h2,
                        eps2=np.array([0.0, 1e-4]), **out)


if __name__ == "__main__":
    main()
