"""Initial conditions in Starlab's conventions (N-body units, G = M = 1).

* ``Ran2`` -- Starlab's uniform generator: ``srandinter``/``randinter`` over
  Numerical Recipes' ``ran2`` (L'Ecuyer's two combined LCGs with a
  Bays-Durham shuffle), src/std/randinter.C:78-95 and :257-324.
* ``makeplummer`` -- the Plummer recipe of src/node/dyn/init/makeplummer.C:55-214
  (Aarseth, Henon & Wielen 1974): shell-stratified mass coordinate with
  mfrac = 0.999, von Neumann rejection for speeds, x/1.695 rescale, centre of
  mass frame, virial ratio 0.5 and E = -1/4 (scale.C:157-190), then the
  random reshuffle of makeplummer.C:236-253.

The potential energy used by the scaling step is O(N^2); callers with large N
pass ``potential_fn`` (e.g. one computed by libg6 on the device).
"""
from __future__ import annotations

import math

import numpy as np

_M1, _M2 = 2147483563, 2147483399
_A1, _A2 = 40014, 40692
_Q1, _Q2 = 53668, 52774
_R1, _R2 = 12211, 3791
_NTAB = 32
_NDIV = 1 + (_M1 - 1) // _NTAB
_RNMX = 1.0 - 1.e-14


def _lcg(s, a, q, r, m):
    k = s // q            # both operands positive: floor == C truncation
    s = a * (s - k * q) - k * r
    if s < 0:
        s += m
    return s


class Ran2:
    """randinter(a, b) = a + (b - a) * ran2(randx) with Starlab's seeding."""

    def __init__(self, seed: int):
        self.seed(seed)

    def seed(self, seed: int) -> int:
        s = abs(int(seed))
        self._state = -s
        self._state2 = 123456789
        self._y = 0
        self._v = [0] * _NTAB
        self.calls = 0
        return s

    def _next(self) -> float:
        if self._state <= 0:
            s = -self._state
            if s < 1:
                s = 1
            self._state2 = s
            for t in range(_NTAB + 7, -1, -1):
                s = _lcg(s, _A1, _Q1, _R1, _M1)
                if t < _NTAB:
                    self._v[t] = s
            self._y = self._v[0]
            self._state = s
        self._state = _lcg(self._state, _A1, _Q1, _R1, _M1)
        self._state2 = _lcg(self._state2, _A2, _Q2, _R2, _M2)
        slot = self._y // _NDIV
        self._y = self._v[slot] - self._state2
        self._v[slot] = self._state
        if self._y < 1:
            self._y += _M1 - 1
        u = (1.0 / _M1) * self._y
        return _RNMX if u > _RNMX else u

    def randinter(self, a: float, b: float) -> float:
        self.calls += 1
        return a + (b - a) * self._next()


def potential_energy(m, x, eps2: float = 0.0, block: int = 2048) -> float:
    """Sum over pairs of -m_i m_j / sqrt(r^2 + eps2) (numpy, O(N^2))."""
    n = len(m)
    tot = 0.0
    for s in range(0, n, block):
        e = min(n, s + block)
        d = x[None, :, :] - x[s:e, None, :]
        r2 = (d * d).sum(-1) + eps2
        with np.errstate(divide="ignore"):
            inv = 1.0 / np.sqrt(r2)
        idx = np.arange(s, e)
        inv[np.arange(e - s), idx] = 0.0
        tot += -(m[s:e, None] * m[None, :] * inv).sum()
    return 0.5 * tot


def makeplummer(n: int, seed: int, scale: bool = True, potential_fn=None):
    """Return (m[n], x[n,3], v[n,3]) as `makeplummer -n n -s seed`."""
    rng = Ran2(seed)
    ri = rng.randinter
    mfrac = 0.999
    two_pi = 2.0 * math.pi
    m = np.full(n, 1.0 / n)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    sqrt2 = math.sqrt(2.0)
    for i in range(n):
        m_min = (i * mfrac) / n
        m_max = ((i + 1) * mfrac) / n
        rrrr = ri(m_min, m_max)
        radius = 1 / math.sqrt(math.pow(rrrr, -2.0 / 3.0) - 1)
        theta = math.acos(ri(-1.0, 1.0))
        phi = ri(0.0, two_pi)
        x[i, 0] = radius * math.sin(theta) * math.cos(phi)
        x[i, 1] = radius * math.sin(theta) * math.sin(phi)
        x[i, 2] = radius * math.cos(theta)
        xx, yy = 0.0, 0.1
        while yy > xx * xx * math.pow(1.0 - xx * xx, 3.5):
            xx = ri(0.0, 1.0)
            yy = ri(0.0, 0.1)
        velocity = xx * sqrt2 * math.pow(1.0 + radius * radius, -0.25)
        theta = math.acos(ri(-1.0, 1.0))
        phi = ri(0.0, two_pi)
        v[i, 0] = velocity * math.sin(theta) * math.cos(phi)
        v[i, 1] = velocity * math.sin(theta) * math.sin(phi)
        v[i, 2] = velocity * math.cos(theta)

    xfac = 1 / 1.695
    vfac = 1 / math.sqrt(xfac)
    x *= xfac
    v *= vfac

    # centre-of-mass frame, accumulated in particle order like the C oracle
    mt = 0.0
    cx = [0.0, 0.0, 0.0]
    cv = [0.0, 0.0, 0.0]
    for i in range(n):
        mt += m[i]
        for k in range(3):
            cx[k] += m[i] * x[i, k]
            cv[k] += m[i] * v[i, k]
    for k in range(3):
        cx[k] /= mt
        cv[k] /= mt
        x[:, k] -= cx[k]
        v[:, k] -= cv[k]

    if scale and n > 1:
        pot = potential_fn(m, x) if potential_fn else potential_energy(m, x)
        kin = 0.5 * float((m * (v * v).sum(1)).sum())
        q = -0.5
        vscale = math.sqrt(q * pot / kin)
        v *= vscale
        kin = q * pot
        energy = kin + pot
        if energy < 0:
            xscale = energy / -0.25
            vs2 = math.sqrt(1.0 / xscale)
            x *= xscale
            v *= vs2

    for i in range(n):
        j = int(ri(0, n))
        if j != i:
            m[[i, j]] = m[[j, i]]
            x[[i, j]] = x[[j, i]]
            v[[i, j]] = v[[j, i]]
    return m, x, v
