"""Batched ensemble engine of libg6.so (include/g6ens.h) and its workloads.

``Ensemble`` mirrors the C entry points (g6e_open/submit/run/poll/result).
The job generators build the two kinds of system the serving workload mixes:

* ``plummer_cluster`` -- a Plummer sphere drawn with the makeplummer recipe
  (src/node/dyn/init/makeplummer.C:55-193: quiet-start mass shells with
  mfrac = 0.999, isotropic directions, von Neumann speeds, x/1.695 rescale,
  centre-of-mass frame) from a numpy generator; like ``makeplummer -u`` it is
  left unscaled (virial units to ~1%), which avoids an O(N^2) rescale per job.
* ``binary_single`` -- a binary-single scattering in the conventions of
  scatter3 (include/scatter3.h:132-164, src/node/dyn/hdyn/sdyn3/evolve/
  scatter3.C:85-110): target binary m1 + m2 = 1, a = 1, thermal
  eccentricity, projectile speed at infinity in units of the critical
  velocity v_c^2 = m1 m2 (m1 + m2 + m3) / m3, impact parameter uniform in
  area up to the focused radius for pericentre <= 4 a, random orientation,
  started at separation r_init on the incoming hyperbola.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import load


def _dp(a):
    return a.ctypes.data_as(C.c_void_p)


class Ensemble:
    """One open engine.  Options are library-wide and outlive g6e_close; this
    wrapper puts every option it changed back to its earlier value on close."""

    def __init__(self, hbm_budget_bytes: int, max_systems: int, **options):
        self.L = load()
        self._saved = {}
        for k, v in options.items():
            self.option(k, v)
        rc = self.L.g6e_open(int(hbm_budget_bytes), int(max_systems))
        if rc:
            self._restore()
            raise RuntimeError(f"g6e_open failed: {self.last_error()}")
        self._n = {}

    def close(self):
        self.L.g6e_close()
        self._restore()

    def _restore(self):
        for k, v in self._saved.items():
            self.L.g6e_set_option(k.encode(), int(v))
        self._saved = {}

    def option(self, key, value=None):
        if value is None:
            return self.L.g6e_get_option(key.encode())
        if key not in self._saved:
            self._saved[key] = self.L.g6e_get_option(key.encode())
        if self.L.g6e_set_option(key.encode(), int(value)):
            raise KeyError(key)
        return value

    def submit(self, m, x, v, t_end, eta, eps2) -> int:
        m = np.ascontiguousarray(m, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        jid = self.L.g6e_submit(len(m), _dp(m), _dp(x), _dp(v), float(t_end), float(eta),
                                float(eps2))
        if jid < 0:
            raise RuntimeError(f"g6e_submit failed: {self.last_error()}")
        self._n[jid] = len(m)
        return jid

    def run(self, steps: int) -> int:
        rc = self.L.g6e_run(int(steps))
        if rc < 0:
            raise RuntimeError(f"g6e_run failed: {self.last_error()}")
        return rc

    def poll(self, max_jobs: int = 65536):
        buf = np.zeros(max_jobs, np.int64)
        k = self.L.g6e_poll(max_jobs, _dp(buf))
        return [int(q) for q in buf[:k]]

    def result(self, jid: int):
        n = self._n.pop(jid)
        x = np.zeros((n, 3)); v = np.zeros((n, 3))
        t = C.c_double(); st = C.c_longlong(); e0 = C.c_double(); e1 = C.c_double()
        rc = self.L.g6e_result(jid, n, _dp(x), _dp(v), C.byref(t), C.byref(st), C.byref(e0),
                               C.byref(e1))
        if rc < 0:
            raise RuntimeError(f"g6e_result failed: {self.last_error()}")
        return dict(x=x, v=v, t=t.value, steps=st.value, e0=e0.value, e1=e1.value,
                    failed=bool(rc))

    def stats(self):
        es = C.c_longlong(); pr = C.c_double(); fin = C.c_longlong(); ds = C.c_double()
        self.L.g6e_stats(C.byref(es), C.byref(pr), C.byref(fin), C.byref(ds))
        return dict(engine_steps=es.value, pairs=pr.value, finished=fin.value,
                    device_seconds=ds.value)

    def last_error(self) -> str:
        s = self.L.g6e_last_error()
        return s.decode() if s else ""

    def run_until_idle(self, chunk: int = 64):
        """Run until every submitted job finished; returns {job: result}."""
        out = {}
        while True:
            left = self.run(chunk)
            for jid in self.poll():
                out[jid] = self.result(jid)
            if left == 0:
                for jid in self.poll():
                    out[jid] = self.result(jid)
                return out


# ---------------------------------------------------------------------------
# shape autotuner
# ---------------------------------------------------------------------------

# Engine options the tuner chooses among: j per split within a system (the
# split-j shape of the step kernels), engine steps per graph replay, and
# whether medium systems run in a workgroup kernel, one workgroup per system
# (k_ens_wg for shared-step jobs, k_ens_wgb for block-step jobs).  The option
# heavy_lanes is not among them: on config 4's clusters two heavy lanes ran
# the whole batch to t = 1 13 % faster, but probes of 1024 to 4096 of the
# systems to t = 1/128 .. 1/8 scored them 0-3 % apart, within the probe's
# noise.  The gain comes once the block schedules have spread and small
# blocks (latency-bound select and correct kernels) dominate, which a short
# probe does not reach.
TUNE_SPACE = dict(jsplit_target=(256, 512), steps_per_graph=(8, 32), wg_steps=(0, 32))
TUNE_SPACE_BLOCK = dict(jsplit_target=(256, 512), steps_per_graph=(8, 32), wgb_steps=(0, 64))


def autotune(jobs, sample=1024, t_probe=1.0 / 128, budget=None, max_systems=8192,
             space=None, fixed=None):
    """Choose the ensemble engine's shape options for a workload.

    Takes a size-stratified sample of `jobs` (dicts with m, x, v, eta, eps2
    and optionally prec and t_end), runs copies of it to t = t_probe (or the
    job's own t_end if earlier) once per candidate option set (the product of
    `space`; default TUNE_SPACE, or TUNE_SPACE_BLOCK when the jobs run on
    block steps) on an idle engine, and returns (best options, {options
    tuple: seconds per pair interaction}, seconds spent).  The jobs
    themselves are not touched.  `fixed` options are applied to every
    candidate (block_steps among them: the jobs' step mode, else the
    engine's current one)."""
    import itertools
    import time as _time
    if space is None:
        block = (fixed or {}).get("block_steps")
        if block is None:
            block = load().g6e_get_option(b"block_steps")
        space = TUNE_SPACE_BLOCK if block else TUNE_SPACE
    order = sorted(range(len(jobs)), key=lambda k: len(jobs[k]["m"]))
    step = max(1, len(order) // max(1, sample))
    probe = [jobs[k] for k in order[::step]][:sample]
    if budget is None:
        # room for the sample twice over (~1.2 KB of pool per particle): opening
        # a pool far larger than the probe needs costs more than the probe
        budget = max(1 << 30, 2 * 1280 * sum(len(j["m"]) for j in probe))
    keys = sorted(space)
    timings = {}
    t_start = _time.perf_counter()
    best, best_cost = None, float("inf")
    for values in itertools.product(*(space[k] for k in keys)):
        opts = dict(zip(keys, values))
        opts.update(fixed or {})
        eng = Ensemble(budget, max_systems, **opts)
        try:
            for j in probe:
                eng.option("precision", j.get("prec", 64))
                eng.submit(j["m"], j["x"], j["v"], min(t_probe, j.get("t_end", t_probe)),
                           j["eta"], j["eps2"])
            t0 = _time.perf_counter()
            eng.run_until_idle(chunk=64)
            wall = _time.perf_counter() - t0
            pairs = eng.stats()["pairs"]
        finally:
            eng.close()
        cost = wall / max(pairs, 1.0)
        timings[values] = cost
        if cost < best_cost:
            best, best_cost = dict(zip(keys, values)), cost
    return best, timings, _time.perf_counter() - t_start


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def _isotropic(rng, n):
    ct = rng.uniform(-1.0, 1.0, n)
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    return np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)


def plummer_cluster(n: int, rng: np.random.Generator):
    """Equal-mass Plummer sphere (makeplummer -u recipe, numpy generator)."""
    mfrac = 0.999
    i = np.arange(n)
    mm = rng.uniform(i * mfrac / n, (i + 1) * mfrac / n)
    r = 1.0 / np.sqrt(mm ** (-2.0 / 3.0) - 1.0)
    x = r[:, None] * _isotropic(rng, n)
    q = np.zeros(n)
    todo = np.arange(n)
    while todo.size:
        xx = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy <= xx * xx * (1.0 - xx * xx) ** 3.5
        q[todo[ok]] = xx[ok]
        todo = todo[~ok]
    speed = q * math.sqrt(2.0) * (1.0 + r * r) ** -0.25
    v = speed[:, None] * _isotropic(rng, n)
    x /= 1.695
    v /= math.sqrt(1 / 1.695)
    m = np.full(n, 1.0 / n)
    x -= (m[:, None] * x).sum(0)
    v -= (m[:, None] * v).sum(0)
    perm = rng.permutation(n)
    return m, x[perm], v[perm]


def _rotation(rng):
    # uniformly random rotation matrix (QR of a Gaussian matrix)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _kepler_state(a, e, mean_anomaly, mtot):
    """Relative position/velocity of a bound orbit in its plane (periapsis on +x)."""
    E = mean_anomaly
    for _ in range(50):
        E = E - (E - e * math.sin(E) - mean_anomaly) / (1 - e * math.cos(E))
    cosE, sinE = math.cos(E), math.sin(E)
    r = a * (1 - e * cosE)
    n = math.sqrt(mtot / a ** 3)
    pos = np.array([a * (cosE - e), a * math.sqrt(1 - e * e) * sinE, 0.0])
    vel = np.array([-a * n * sinE / (1 - e * cosE),
                    a * n * math.sqrt(1 - e * e) * cosE / (1 - e * cosE), 0.0])
    return pos, vel, r


def binary_single(rng: np.random.Generator, m2: float = 0.5, m3: float = 0.5,
                  sigma: float = 0.5, r_init: float = 20.0, r_peri_max: float = 4.0):
    """Binary (m1 = 1 - m2, m2; a = 1) plus projectile m3, centre-of-mass frame."""
    m1 = 1.0 - m2
    mbin = 1.0
    mtot = mbin + m3
    ecc = math.sqrt(rng.uniform())
    pos_b, vel_b, _ = _kepler_state(1.0, ecc, rng.uniform(0, 2 * math.pi), mbin)
    R = _rotation(rng)
    pos_b, vel_b = R @ pos_b, R @ vel_b
    v_inf = sigma * float(np.linalg.norm(rng.normal(size=3)))       # Maxwellian, units of v_c
    v_crit = math.sqrt(m1 * m2 * mtot / m3)
    vi = max(v_inf * v_crit, 1e-3)
    b_max = r_peri_max * math.sqrt(1.0 + 2.0 * mtot / (r_peri_max * vi * vi))
    rho = b_max * math.sqrt(rng.uniform())
    L = rho * vi
    v_r = math.sqrt(max(vi * vi + 2.0 * mtot / r_init - (L / r_init) ** 2, 0.0))
    rel_x = np.array([r_init, 0.0, 0.0])
    rel_v = np.array([-v_r, L / r_init, 0.0])
    x = np.zeros((3, 3))
    v = np.zeros((3, 3))
    x[0] = -m2 / mbin * pos_b
    x[1] = m1 / mbin * pos_b
    v[0] = -m2 / mbin * vel_b
    v[1] = m1 / mbin * vel_b
    x[:2] -= m3 / mtot * rel_x
    v[:2] -= m3 / mtot * rel_v
    x[2] = mbin / mtot * rel_x
    v[2] = mbin / mtot * rel_v
    m = np.array([m1, m2, m3])
    return m, x, v
