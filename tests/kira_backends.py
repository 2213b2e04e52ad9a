"""Force backends for the block-step driver used by the config-1 tests.

HostBackend   -- kira's host path (config 0): every j predicted on the host
                 with predict_loworder (include/_dyn_.h:290-295), FP64 direct
                 sum (the oracle).  This is the all-CPU reference run.
CheckedG6     -- libg6 through the ABI, plus a host mirror of j-memory; after
                 each force call the oracle recomputes the same block from the
                 same inputs (GRAPE-form prediction, bitwise the device's) and
                 the per-block errors are recorded.
OracleAbiBackend -- the arithmetic of tests/oracle_g6abi.c (a GRAPE-6 ABI
                 library on the oracle) in Python, for bitwise comparisons of
                 the Python and native drivers without a GPU.
NativeChecker -- the CheckedG6 checks as an observer of the native driver.
"""
import numpy as np

import oracle
from starlab_amd.blockstep import ONE3, G6Backend
from tests.gpu_common import per_particle, percentiles


class HostBackend:
    def open(self, n):
        self.n = n
        self.x = np.zeros((n, 3)); self.v = np.zeros((n, 3))
        self.a = np.zeros((n, 3)); self.j = np.zeros((n, 3))
        self.t = np.zeros(n); self.m = np.zeros(n)
        self.idx = np.arange(n, dtype=np.int32)

    def send_j(self, idx, tj, dtj, m, x, v, a, j, k18):
        self.t[idx] = tj; self.m[idx] = m; self.x[idx] = x; self.v[idx] = v
        self.a[idx] = a; self.j[idx] = j

    def force(self, nj, t, idx, xi, vi, eps2):
        dt = (t - self.t)[:, None]
        dt3 = dt * ONE3
        dt2 = dt * 0.5
        xp = ((self.j * dt3 + self.a) * dt2 + self.v) * dt + self.x
        vp = (self.j * dt2 + self.a) * dt + self.v
        acc, jerk, pot, nn, _ = oracle.force(xi, vi, idx, xp, vp, self.m, self.idx, eps2)
        return acc, jerk, pot, nn

    def close(self):
        pass


class OracleAbiBackend:
    """The arithmetic of tests/oracle_g6abi.c in Python: j-memory holds what
    g6_set_j_particle_ receives (a/2, j/6, k/18), every j is predicted with
    the oracle's GRAPE-form predictor (4th order) at t_i, and the FP64 oracle
    sums the forces.  The Python driver on this backend and the native driver
    on that library must agree bit for bit."""

    def open(self, n):
        self.n = n
        self.mir = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), a2=np.zeros((n, 3)),
                        j6=np.zeros((n, 3)), k18=np.zeros((n, 3)), t=np.zeros(n), m=np.zeros(n))
        self.idx = np.arange(n, dtype=np.int32)

    def send_j(self, idx, tj, dtj, m, x, v, a, j, k18):
        M = self.mir
        M["t"][idx] = tj; M["m"][idx] = m; M["x"][idx] = x; M["v"][idx] = v
        M["a2"][idx] = a / 2.0; M["j6"][idx] = j / 6.0; M["k18"][idx] = k18

    def force(self, nj, t, idx, xi, vi, eps2):
        M = self.mir
        xp, vp = oracle.predict(M["x"], M["v"], M["a2"], M["j6"], M["k18"], M["t"], t, 4)
        acc, jerk, pot, nn, _ = oracle.force(xi, vi, idx, xp, vp, M["m"], self.idx, eps2)
        return acc, jerk, pot, nn

    def close(self):
        pass


class _ErrorBook:
    """Per-block errors of library results against the FP64 oracle on the
    same inputs: normwise acc error |da| / sum_j |a_ij|, jerk likewise
    (|dj| / sum_j |j_ij| bound), pot relative, nearest-neighbour mismatches."""

    def _book_init(self, every, max_i, seed):
        self.worst = dict(acc_norm=0.0, jerk_norm=0.0, pot_rel=0.0, acc_rel=0.0)
        self.acc_errs = []
        self.rel_errs = dict(acc=[], jerk=[], pot=[])
        self.nn_mismatch = 0
        self.checked_blocks = 0
        self.checked_i = 0
        self.calls = 0
        self.every = every
        self.max_i = max_i
        self.rng = np.random.default_rng(seed)

    def _book(self, xp, vp, mass, jidx, t, idx, xi, vi, acc, jerk, pot, nn, eps2):
        """Check one force call (if it is an `every`-th one); xp, vp = the
        oracle's prediction of all j at t, or a callable producing them."""
        self.calls += 1
        if (self.calls - 1) % self.every:
            return
        sel = np.arange(len(idx))
        if self.max_i is not None and len(idx) > self.max_i:
            sel = np.sort(self.rng.choice(len(idx), self.max_i, replace=False))
        acc, jerk, pot, nn = acc[sel], jerk[sel], pot[sel], nn[sel]
        idx, xi, vi = idx[sel], xi[sel], vi[sel]
        if callable(xp):
            xp, vp = xp()
        racc, rjerk, rpot, rnn, _ = oracle.force(xi, vi, idx, xp, vp, mass, jidx, eps2)
        sa = np.zeros(len(idx)); sj = np.zeros(len(idx))
        for k, i in enumerate(idx):
            d = xp - xi[k]; w = vp - vi[k]
            r2 = (d * d).sum(1) + eps2
            r2[i] = np.inf
            sa[k] = (mass / r2).sum()
            sj[k] = (mass * 4.0 * np.sqrt((w * w).sum(1)) / (r2 * np.sqrt(r2))).sum()
        ea = np.linalg.norm(acc - racc, axis=1) / sa
        self.acc_errs.append(ea)
        ra, rj, rp = per_particle(acc, jerk, pot, racc, rjerk, rpot)
        self.rel_errs["acc"].append(ra)
        self.rel_errs["jerk"].append(rj)
        self.rel_errs["pot"].append(rp)
        self.worst["acc_rel"] = max(self.worst["acc_rel"], float(ra.max()))
        self.worst["acc_norm"] = max(self.worst["acc_norm"], float(ea.max()))
        self.worst["jerk_norm"] = max(self.worst["jerk_norm"],
                                      float((np.linalg.norm(jerk - rjerk, axis=1) / sj).max()))
        self.worst["pot_rel"] = max(self.worst["pot_rel"],
                                    float((np.abs(pot - rpot) / np.abs(rpot)).max()))
        self.nn_mismatch += int((nn != rnn).sum())
        self.checked_blocks += 1
        self.checked_i += len(idx)

    def summary(self):
        """Per-particle relative error percentiles (median, p99, max) of acc,
        jerk and pot, and the normwise acc percentiles, over all checks."""
        out = {k: percentiles(np.concatenate(v) if v else []) for k, v in self.rel_errs.items()}
        norm = np.concatenate(self.acc_errs) if self.acc_errs else np.zeros(0)
        out["acc_norm"] = percentiles(norm)
        out["acc_norm"]["p999"] = float(np.percentile(norm, 99.9)) if norm.size else 0.0
        out["jerk_norm_max"] = self.worst["jerk_norm"]
        out["checked_particle_steps"] = self.checked_i
        return out


class CheckedG6(G6Backend, _ErrorBook):
    """libg6 through the ABI with a host mirror of j-memory.  every: check
    every `every`-th force call; max_i: at most this many i of a checked call
    (a seeded sample) go to the oracle."""

    def __init__(self, g, bulk=False, every=1, max_i=None, seed=0):
        super().__init__(g, bulk)
        self._book_init(every, max_i, seed)

    def open(self, n):
        super().open(n)
        self.mir = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), a2=np.zeros((n, 3)),
                        j6=np.zeros((n, 3)), k18=np.zeros((n, 3)), t=np.zeros(n), m=np.zeros(n))
        self.idx = np.arange(n, dtype=np.int32)

    def send_j(self, idx, tj, dtj, m, x, v, a, j, k18):
        super().send_j(idx, tj, dtj, m, x, v, a, j, k18)
        M = self.mir
        M["t"][idx] = tj; M["m"][idx] = m; M["x"][idx] = x; M["v"][idx] = v
        M["a2"][idx] = a / 2.0; M["j6"][idx] = j / 6.0; M["k18"][idx] = k18

    def force(self, nj, t, idx, xi, vi, eps2):
        acc, jerk, pot, nn = super().force(nj, t, idx, xi, vi, eps2)
        M = self.mir
        pred = lambda: oracle.predict(M["x"], M["v"], M["a2"], M["j6"], M["k18"], M["t"], t, 4)
        self._book(pred, None, M["m"], self.idx, t, idx, xi, vi, acc, jerk, pot, nn, eps2)
        return acc, jerk, pot, nn


class NativeChecker(_ErrorBook):
    """Observer for NativeBlockHermite: the same oracle checks as CheckedG6,
    from the j-memory state the native driver reports at each force call."""

    def __init__(self, eps2, every=1, max_i=None, seed=0):
        self.eps2 = eps2
        self._book_init(every, max_i, seed)

    def __call__(self, jm, t, idx, xi, vi, acc, jerk, pot, nn):
        pred = lambda: oracle.predict(jm["x"], jm["v"], jm["acc"] / 2.0, jm["jerk"] / 6.0,
                                      jm["k18"], jm["t"], t, 4)
        jidx = np.arange(len(jm["m"]), dtype=np.int32)
        self._book(pred, None, jm["m"].copy(), jidx, t, idx.copy(), xi.copy(), vi.copy(),
                   acc.copy(), jerk.copy(), pot.copy(), nn.copy(), self.eps2)
        return False


class CallMixChecker(_ErrorBook):
    """Observer for KiraCallMix (bs_run_kira): every `every`-th force call of
    any kind (blocks, single-node redos, perturbation pairs, energy sweeps,
    density rounds) against the FP64 oracle on the j-memory exactly as sent;
    every `list_every`-th neighbour list against the brute-force FP64 set with
    the library's conventions (the i itself included when npipes > 48,
    overflow iff the set has max_length or more entries, then the first
    max_length addresses).  Counts calls by kind, list reads by status, and
    calls whose return code or nn flagged an (injected) error."""

    def __init__(self, eps2, npipes, every=1, max_i=None, seed=0, list_every=1):
        from collections import Counter
        self.eps2 = eps2
        self.include_self = npipes > 48
        self.kinds = Counter()
        self.read_status = Counter()
        self.errors = 0
        self.list_checked = 0
        self.list_mismatch = []
        self.list_every = list_every
        self.lists_seen = 0
        self._book_init(every, max_i, seed)

    def _pred(self, jm, ti):
        return oracle.predict(jm["x"], jm["v"], jm["a2"], jm["j6"], jm["k18"], jm["t"], ti, 4)

    def __call__(self, jm, c):
        from starlab_amd.blockstep import CALL_LIST, NOT_CALLED
        if c["read_status"] != NOT_CALLED:
            self.read_status[c["read_status"]] += 1
            return False
        self.kinds[c["kind"]] += 1
        n = len(jm["m"])
        jidx = np.arange(n, dtype=np.int32)
        if c["kind"] == CALL_LIST:
            self.lists_seen += 1
            if (self.lists_seen - 1) % self.list_every:
                return False
            xp, _ = self._pred(jm, c["ti"])
            mx = c["max_length"]
            counts, lists = oracle.neighbours(c["xi"].copy(), c["h2"].copy(), c["index"].copy(),
                                              self.include_self, xp, jidx, mx)
            cnt = int(counts[0])
            want_rc = 1 if cnt >= mx else 0
            want = lists[0][:min(cnt, mx)]
            got = c["nbl"]
            self.list_checked += 1
            if c["rc"] != want_rc or len(got) != len(want) or not np.array_equal(got, want):
                self.list_mismatch.append((c["ti"], int(c["index"][0]), cnt, c["rc"], len(got)))
            return False
        if c["rc"] != 0 or np.any(c["nn"] < 0):
            self.errors += 1
            return False
        self._book(lambda: self._pred(jm, c["ti"]), None, jm["m"].copy(), jidx, c["ti"],
                   c["index"].copy(), c["xi"].copy(), c["vi"].copy(), c["acc"].copy(),
                   c["jerk"].copy(), c["pot"].copy(), c["nn"].copy(), self.eps2)
        return False
