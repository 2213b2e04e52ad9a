"""Shared helpers for the GPU parity tests (inputs, oracle comparison)."""
from __future__ import annotations

import numpy as np

import oracle

# Tolerances (FP32 pair arithmetic on double-single positions, FP64 sums,
# the nearest-neighbour tile recomputed in FP64):
#   pot:   max |dpot/pot|                      <= POT_TOL
#   acc:   max |da| / sum_j |a_ij|             <= ACC_TOL   (normwise)
#   jerk:  max |dj| / sum_j |j_ij|             <= JERK_TOL  (normwise)
#   acc:   max |da| / |a|                      <= ACC_REL_TOL (per particle;
#          asserted where no particle's acceleration cancels to ~0)
#   nn:    exact index agreement
# Per-particle jerk error |dj| / |j| is reported, not bounded: where the
# pair terms cancel (|j| << sum_j |j_ij|) FP32 pair arithmetic cannot hold it
# at 1e-6.
POT_TOL = 1e-6
ACC_TOL = 1e-6
JERK_TOL = 1e-6
ACC_REL_TOL = 1e-6


def percentiles(a):
    """median, p99, max of an error array (a dict of floats)."""
    a = np.asarray(a, dtype=np.float64)
    if a.size == 0:
        return dict(median=0.0, p99=0.0, max=0.0)
    return dict(median=float(np.median(a)), p99=float(np.percentile(a, 99)),
                max=float(a.max()))


def per_particle(acc, jerk, pot, racc, rjerk, rpot):
    """Per-particle relative errors |da|/|a|, |dj|/|j|, |dpot/pot|."""
    ea = np.linalg.norm(acc - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ej = np.linalg.norm(jerk - rjerk, axis=1) / np.linalg.norm(rjerk, axis=1)
    ep = np.abs(pot - rpot) / np.abs(rpot)
    return ea, ej, ep


def plummer_state(n, seed=42, tspread=2.0 ** -10, k18=False):
    """Plummer m, x, v plus synthetic predictor inputs (a2, j6, k18, tj)."""
    m, x, v = oracle.makeplummer(n, seed)
    rng = np.random.default_rng(seed)
    idx = np.arange(n, dtype=np.int32)
    acc, jerk, _, _, _ = oracle.force(x, v, idx, x, v, m, idx, 1e-8) if n <= 8192 else (
        rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), None, None, None)
    snap = rng.normal(scale=10.0, size=(n, 3)) if k18 else np.zeros((n, 3))
    tj = -rng.uniform(0.0, tspread, n)
    return dict(m=m, x=x, v=v, a2=acc / 2, j6=jerk / 6, k18=snap / 18, tj=tj,
                dtj=np.full(n, 2.0 ** -10), idx=idx)


def load_j(g, s, bulk=True):
    n = len(s["m"])
    if bulk:
        rc = g.set_j_particles(s["idx"], s["idx"], s["tj"], s["dtj"], s["m"], s["k18"],
                               s["j6"], s["a2"], s["v"], s["x"])
        assert rc == 0
    else:
        for a in range(n):
            rc = g.set_j_particle(a, int(s["idx"][a]), s["tj"][a], s["dtj"][a], s["m"][a],
                                  s["k18"][a], s["j6"][a], s["a2"][a], s["v"][a], s["x"][a])
            assert rc == 0


def term_scales(xi, mj, xj, eps2, excl):
    """sum_j m_j/r^2 (acc) and sum_j m_j |w|/r^3 + 3 m_j |w|/r^3 bound (jerk)."""
    sa = np.zeros(len(xi))
    for k in range(len(xi)):
        d = xj - xi[k]
        r2 = (d * d).sum(1) + eps2
        r2[excl[k]] = np.inf if 0 <= excl[k] < len(xj) else r2[excl[k]]
        sa[k] = (mj / r2).sum()
    return sa


def compare(acc, jerk, pot, nn, ref, xi, vi, mj, xj, vj, eps2, excl):
    """Return error summary dict; ref = oracle.force(...) tuple."""
    racc, rjerk, rpot, rnn, _ = ref
    n = len(xi)
    sa = np.zeros(n)
    sj = np.zeros(n)
    for k in range(n):
        d = xj - xi[k]
        w = vj - vi[k]
        r2 = (d * d).sum(1) + eps2
        if 0 <= excl[k] < len(xj):
            r2[excl[k]] = np.inf
        r = np.sqrt(r2)
        sa[k] = (mj / r2).sum()
        sj[k] = (mj * (np.sqrt((w * w).sum(1)) * 4.0) / (r2 * r)).sum()
    ea = np.linalg.norm(acc - racc, axis=1)
    ej = np.linalg.norm(jerk - rjerk, axis=1)
    ra, rj, rp = per_particle(acc, jerk, pot, racc, rjerk, rpot)
    pa, pj, pp = percentiles(ra), percentiles(rj), percentiles(rp)
    out = dict(
        acc_norm=float((ea / sa).max()),
        jerk_norm=float((ej / sj).max()),
        pot_rel=pp["max"],
        acc_rel_median=pa["median"], acc_rel_p99=pa["p99"], acc_rel_max=pa["max"],
        jerk_rel_median=pj["median"], jerk_rel_p99=pj["p99"], jerk_rel_max=pj["max"],
        pot_rel_median=pp["median"], pot_rel_p99=pp["p99"],
        nn_mismatch=int((nn != rnn).sum()),
    )
    return out
