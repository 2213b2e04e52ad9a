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


# ---------------------------------------------------------------------------
# King models: src/node/dyn/init/makeking.C (King 1966; Steve McMillan and
# Kimberly Engle).  The model is the solution of (x psi)'' = 9 x d(psi) from
# the centre (psi = -W0, d = 1) to the tidal radius (d = 0), with
# x = r / rc, psi = phi / sig^2, d = rho / rho0 and the lowered-Maxwellian
# distribution f ~ exp(-psi) (exp(-v^2 / 2 sig^2) - exp(psi - beta W0)).
# ---------------------------------------------------------------------------
_K_NM, _K_NINDX, _K_NG, _K_YMAX = 2500, 100, 1000, 4.0
_K_RLIN, _K_NLIN, _K_RMAX = 0.25, 105, 1e4


def _trapzd(f, a, b, eps=1e-10, max_steps=50):
    """Trapezoid refinement until successive sums agree to eps (relative),
    makeking.C:101-150 (Numerical Recipes' trapzd driven to convergence)."""
    s = 0.0
    prev = -1.0e30
    for n in range(1, max_steps + 1):
        if n == 1:
            s = 0.5 * (b - a) * (f(a) + f(b))
        else:
            it = 1 << (n - 2)
            dx = (b - a) / it
            xs = a + 0.5 * dx + dx * np.arange(it)
            s = 0.5 * (s + (b - a) / it * float(np.sum(f(xs))))
        if abs(s - prev) < eps * abs(prev):
            return s
        prev = s
    return 0.0


class KingModel:
    """Profile of a King model of dimensionless depth W0 (makeking.C:152-472)
    and the quantities makeking derives from it (makeking.C:687-757)."""

    def __init__(self, w0: float, beta: float = 0.0):
        if w0 > 16:
            raise ValueError("makeking: W0 < 16 required")
        if beta >= 1:
            raise ValueError("makeking: beta < 1 required")
        self.w0 = w0
        self.beta = beta
        self.beta_w0 = beta * w0
        self.scale_fac = math.exp(self.beta_w0)
        self._g_table()
        self._poisson()
        self._statistics()

    # g[i] = integral_0^{i YMAX/NG} y^2 exp(-y^2) dy   (makeking.C:160-172)
    def _g_table(self):
        gaus2 = lambda y: y * y * np.exp(-y * y)
        g = np.zeros(_K_NG + 1)
        dy = _K_YMAX / _K_NG
        yo = 0.0
        for i in range(1, _K_NG + 1):
            y = yo + dy
            g[i] = g[i - 1] + _trapzd(gaus2, yo, y)
            yo = y
        self.g = g

    def _gg(self, y, want_g4):
        """(g2, g4) = integrals of y^2 e^-y^2 and y^4 e^-y^2 from 0 to y."""
        g = self.g
        yindx = _K_NG * y / _K_YMAX
        iy = int(yindx)
        g2 = g[_K_NG] if iy >= _K_NG else g[iy] + (yindx - iy) * (g[iy + 1] - g[iy])
        g4 = 0.0
        if want_g4:
            y2 = y * y
            g4 = 1.5 * g2 - 0.5 * y * y2 * math.exp(-y2)
        return g2, g4

    def dens_and_vel(self, psi, v2_in=1.0):
        """Density (unnormalised) and local velocity dispersion at scaled
        potential psi < 0 (get_dens_and_vel, makeking.C:197-236)."""
        if psi >= -self.beta_w0:
            return 0.0, v2_in
        p_max = -psi - self.beta_w0
        y_max = math.sqrt(p_max)
        ep = math.exp(-psi)
        g2, g4 = self._gg(y_max, v2_in > 0)
        dens = ep * g2 - self.scale_fac * y_max * p_max / 3
        v2 = v2_in
        if v2_in > 0 and dens > 0:
            v2 = 2 * (ep * g4 - 0.2 * self.scale_fac * y_max * p_max * p_max) / dens
        return dens, v2

    def _rhs(self, y, x):
        if x <= 0:
            d = 1.0
        else:
            d, _ = self.dens_and_vel(y[0] / x, 0.0)
            d = d * self._dc_inverse
        return (y[1], 9 * x * d)

    def _rk4(self, x, x_next, y, dx):
        """Classical RK4 steps of dx until x reaches x_next (makeking.C:243-304)."""
        while x < x_next - 0.5 * dx:
            d0 = self._rhs(y, x)
            dy1 = (dx * d0[0], dx * d0[1])
            y1 = (y[0] + 0.5 * dy1[0], y[1] + 0.5 * dy1[1])
            d1 = self._rhs(y1, x + 0.5 * dx)
            dy2 = (dx * d1[0], dx * d1[1])
            y2 = (y[0] + 0.5 * dy2[0], y[1] + 0.5 * dy2[1])
            d2 = self._rhs(y2, x + 0.5 * dx)
            dy3 = (dx * d2[0], dx * d2[1])
            y3 = (y[0] + dy3[0], y[1] + dy3[1])
            d3 = self._rhs(y3, x + dx)
            dy4 = (dx * d3[0], dx * d3[1])
            y = (y[0] + (dy1[0] + 2 * dy2[0] + 2 * dy3[0] + dy4[0]) / 6.0,
                 y[1] + (dy1[1] + 2 * dy2[1] + 2 * dy3[1] + dy4[1]) / 6.0)
            x += dx
        return x, y

    def _poisson(self):
        """Integrate out to the tidal radius (poisson, makeking.C:310-472):
        RLIN core radii on a linear grid of NLIN points, then logarithmic out
        to RMAX, about 20 RK4 steps per grid interval."""
        nmax = _K_NM
        rr = np.zeros(nmax + 1); d = np.zeros(nmax + 1); v2 = np.zeros(nmax + 1)
        psi = np.zeros(nmax + 1); zm = np.zeros(nmax + 1)
        psi0 = -abs(self.w0)
        xn = 0.0
        y = (0.0, psi0)
        psi[0] = psi0
        v2[0] = 1.0
        d[0], v2[0] = self.dens_and_vel(psi0, 1.0)
        self._dc_inverse = 1.0 / d[0]
        fac = math.pow(10, math.log10(_K_RMAX / _K_RLIN) / (nmax - _K_NLIN))
        i = 0
        done = False
        for i in range(1, nmax + 1):
            xo = xn
            xn = (_K_RLIN * i) / _K_NLIN if i <= _K_NLIN else fac * xo
            dx = 0.051 * (xn - xo)
            xo, y = self._rk4(xo, xn, y, dx)
            xn = xo
            rr[i] = xn
            psi[i] = y[0] / xn
            d[i], v2[i] = self.dens_and_vel(psi[i], 1.0)
            if d[i] < 0:
                # interpolate to the tidal radius
                rr[i] = rr[i - 1] + (rr[i] - rr[i - 1]) / (0.1 - d[i] / d[i - 1])
                d[i] = 0.0
                v2[i] = 0.0
            zm[i] = rr[i] * y[1] - y[0]
            if not d[i] > 0:
                done = True
                break
        nprof = i if done else nmax
        self.v20 = v2[0]
        d = d / d[0]
        v2 = v2 / v2[0]
        zm = (4.0 * math.pi / 9) * zm
        self.nprof = nprof
        self.rr, self.d, self.v2, self.psi, self.zm = (
            rr[:nprof + 1], d[:nprof + 1], v2[:nprof + 1], psi[:nprof + 1], zm[:nprof + 1])

    def _statistics(self):
        """Mass index, core mass, half-mass, virial radius (makeking.C:696-757)."""
        nprof = self.nprof
        rr = self.rr
        self.rho0 = 1.0 / self.zm[nprof]
        self.sig = math.sqrt(4.0 * math.pi / 3.0 * self.rho0 / 3)
        zm = self.zm / self.zm[nprof]
        self.zm = zm
        indx = np.zeros(_K_NINDX + 1, dtype=np.int64)
        indx[_K_NINDX] = nprof
        dz = 1.0 / _K_NINDX
        z = dz
        iz = 1
        jcore = jhalf = 0
        for j in range(1, nprof):
            if rr[j] < 1:
                jcore = j
            if zm[j] < 0.5:
                jhalf = j
            if zm[j] > z:
                indx[iz] = j - 1
                z = z + dz
                iz += 1
        self.indx = indx
        self.zmcore = zm[jcore] + (zm[jcore + 1] - zm[jcore]) * (1 - rr[jcore]) / (
            rr[jcore + 1] - rr[jcore])
        self.rhalf = rr[jhalf] + (rr[jhalf + 1] - rr[jhalf]) * (0.5 - zm[jhalf]) / (
            zm[jhalf + 1] - zm[jhalf])
        pot = 0.0
        for i in range(1, nprof + 1):
            pot -= (zm[i] - zm[i - 1]) * (zm[i] + zm[i - 1]) / (rr[i - 1] + rr[i])
        self.rvirial = -0.5 / pot
        self.rtidal = rr[nprof]
        self.concentration = math.log10(rr[nprof])
        step = _K_YMAX / _K_NG
        self.v33 = self.scale_fac * (step * np.arange(_K_NG + 1)) ** 3 / 3

    # ---- sampling (setpos / setvel, makeking.C:514-642) ----------------------
    def position(self, ri):
        """Random position from the mass profile; returns (pos, psi)."""
        rno = ri(0.0, 1.0)
        i = int(_K_NINDX * rno)
        zm, rr, psi, indx = self.zm, self.rr, self.psi, self.indx
        i1 = int(indx[i])
        hi = int(indx[i + 1]) + 1
        while i1 <= hi and not zm[i1] > rno:
            i1 += 1
        if i1 > hi:
            raise RuntimeError("makeking: error in getpos")
        rfac = (rno - zm[i1 - 1]) / (zm[i1] - zm[i1 - 1])
        r = rr[i1 - 1] + rfac * (rr[i1] - rr[i1 - 1])
        p = psi[i1 - 1] + rfac * (psi[i1] - psi[i1 - 1])
        cth = 2 * ri(0.0, 1.0) - 1
        sth = math.sqrt(1 - cth * cth)
        ph = 2.0 * math.pi * ri(0.0, 1.0)
        return (r * sth * math.cos(ph), r * sth * math.sin(ph), r * cth), p

    def velocity(self, ri, p):
        """Random velocity at scaled potential p (units of sig), by bisection
        on the cumulative speed distribution exp(-p) g(y) - v33(y)."""
        v = 0.0
        if p < -self.beta_w0:
            g, v33 = self.g, self.v33
            pfac = math.exp(-p)
            il = 0
            iu = int((_K_NG / _K_YMAX) * math.sqrt(-p))
            if iu > _K_NG:
                iu = _K_NG
            rl = 0.0
            ru = pfac * g[iu] - v33[iu]
            rno = ri(0.0, 1.0) * ru
            while iu - il > 1:
                im = (il + iu) // 2
                rm = pfac * g[im] - v33[im]
                if rm > rno:
                    iu, ru = im, rm
                else:
                    il, rl = im, rm
            v = (_K_YMAX / _K_NG) * math.sqrt(2.0) * (il + (rno - rl) / (ru - rl))
        cth = 2 * ri(0.0, 1.0) - 1
        sth = math.sqrt(1 - cth * cth)
        ph = 2.0 * math.pi * ri(0.0, 1.0)
        return (v * sth * math.cos(ph), v * sth * math.sin(ph), v * cth)


def makeking(n: int, w0: float, seed: int, beta: float = 0.0):
    """`makeking -n n -w w0 -s seed -u` (makeking.C:659-842): equal masses
    1/n, virial radius 1, centre-of-mass frame, not scaled to E = -1/4.
    Returns (m, x, v, model)."""
    model = KingModel(w0, beta)
    rng = Ran2(seed)
    ri = rng.randinter
    m = np.full(n, 1.0 / n)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    for i in range(n):
        pos, p = model.position(ri)
        x[i] = pos
        v[i] = model.velocity(ri, p)
    v *= model.sig
    xfac = 1.0 / model.rvirial
    x *= xfac
    v *= 1.0 / math.sqrt(xfac)
    x -= (m[:, None] * x).sum(0) / m.sum()
    v -= (m[:, None] * v).sum(0) / m.sum()
    return m, x, v, model


def power_law_masses(n: int, seed: int, lower: float, upper: float, exponent: float):
    """`makemass -f 1 -x exponent -l lower -u upper -s seed`: dN/dm ~ m^exponent
    on [lower, upper] by inverting the cumulative distribution
    (general_power_law, src/node/util/makemass.C:374-391)."""
    rng = Ran2(seed)
    m = np.empty(n)
    if exponent == -1:
        norm = math.log(upper / lower)
        for i in range(n):
            m[i] = lower * math.exp(norm * rng.randinter(0, 1))
    else:
        norm = math.pow(upper / lower, 1 + exponent) - 1
        inv = 1 / (1 + exponent)
        for i in range(n):
            m[i] = lower * math.pow(norm * rng.randinter(0, 1) + 1, inv)
    return m


def add_secondaries(m, fraction: float, seed: int):
    """`makesecondary -f fraction -s seed` on single top-level stars
    (src/node/util/makesecondary.C:214-360, default branch): every
    1/fraction-th star (running sum of the fraction crossing 1) becomes the
    primary of a binary whose secondary mass is uniform on
    [min mass, primary mass].  Returns (primary index array, secondary
    masses)."""
    rng = Ran2(seed)
    m = np.asarray(m, dtype=np.float64)
    mmin, mmax = float(m.min()), float(m.max())
    prim, msec = [], []
    s = 0.0
    for i, mp in enumerate(m):
        s += fraction
        if s >= 0.9999999999:
            s -= 1
            qmin = mmin / mp
            qmax = min(mmax, mp) / mp
            q = rng.randinter(qmin, qmax)
            prim.append(i)
            msec.append(q * mp)
    return np.array(prim, dtype=np.int64), np.array(msec)


def binary_orbits(m1, m2, seed: int, a_min: float, a_max: float, emax: float = 1.0):
    """Relative orbits as `makebinary -f 2 -o 1 -l a_min -u a_max -e emax`
    (src/node/dyn/init/makebinary.C:184-330): thermal eccentricity
    e = sqrt(U(0, emax)), semi-major axis log-uniform on [a_min, a_max]
    (here in N-body units), isotropic orientation and uniform mean anomaly.
    Returns (a, e, dx, dv): the secondary's position and velocity relative to
    the primary."""
    rng = Ran2(seed)
    ri = rng.randinter
    k = len(m1)
    a = np.empty(k); e = np.empty(k)
    dx = np.zeros((k, 3)); dv = np.zeros((k, 3))
    for b in range(k):
        ecc = math.sqrt(ri(0, emax))
        sma = a_min * math.exp(ri(0.0, math.log(a_max) - math.log(a_min)))
        mt = m1[b] + m2[b]
        # orientation: random unit vectors for the orbit normal and periastron
        cth = ri(-1.0, 1.0); ph = ri(0.0, 2 * math.pi); psi = ri(0.0, 2 * math.pi)
        sth = math.sqrt(1 - cth * cth)
        nrm = np.array([sth * math.cos(ph), sth * math.sin(ph), cth])
        t1 = np.cross(nrm, [0.0, 0.0, 1.0] if abs(cth) < 0.9 else [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(nrm, t1)
        ehat = math.cos(psi) * t1 + math.sin(psi) * t2
        qhat = np.cross(nrm, ehat)
        # mean anomaly -> eccentric anomaly (Newton on Kepler's equation)
        M = ri(0.0, 2 * math.pi)
        E = M if ecc < 0.8 else math.pi
        for _ in range(100):
            f = E - ecc * math.sin(E) - M
            E -= f / (1 - ecc * math.cos(E))
            if abs(f) < 1e-15:
                break
        cE, sE = math.cos(E), math.sin(E)
        rs = math.sqrt(max(1 - ecc * ecc, 0.0))
        r = sma * (1 - ecc * cE)
        n_mean = math.sqrt(mt / sma ** 3)
        pos = sma * (cE - ecc) * ehat + sma * rs * sE * qhat
        vel = (n_mean * sma * sma / r) * (-sE * ehat + rs * cE * qhat)
        a[b], e[b] = sma, ecc
        dx[b], dv[b] = pos, vel
    return a, e, dx, dv


def scale_nbody(m, x, v, potential_fn=None, mass=1.0, q=0.5, energy=-0.25):
    """`scale -m mass -q q -e energy` on top-level nodes
    (src/node/dyn/util/scale.C:81-86, 157-190, 350-470): total mass, then
    the virial ratio by scaling velocities, then the energy by scaling
    positions and velocities at fixed virial ratio (eps = 0)."""
    m = np.asarray(m, dtype=np.float64) * (mass / float(np.sum(m)))
    x = np.array(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    x -= (m[:, None] * x).sum(0) / mass
    v -= (m[:, None] * v).sum(0) / mass
    pot = potential_fn(m, x) if potential_fn else potential_energy(m, x)
    kin = 0.5 * float((m * (v * v).sum(1)).sum())
    qq = -abs(q)
    v *= math.sqrt(qq * pot / kin)
    kin = qq * pot
    e = kin + pot
    if e < 0:
        xscale = e / -abs(energy)
        x *= xscale
        v *= math.sqrt(1.0 / xscale)
    return m, x, v


def king_cluster(n: int, seed: int, w0: float = 6.0, mass_ratio: float = 100.0,
                 exponent: float = -2.35, bin_frac: float = 0.05, a_range=(1e-5, 1e-3),
                 potential_fn=None):
    """The King W0 / Salpeter / primordial-binary cluster:

        makeking -n n -w w0 -u | makemass -f 1 -x exponent -l 1 -u mass_ratio
          | makesecondary -f bin_frac | scale -m 1 -q 0.5 -e -0.25
          | makebinary -f 2 -o 1 -l a_min -u a_max

    with seeds seed, seed + 1, seed + 2, seed + 3.  Returns a dict with the
    top-level nodes (m, x, v: singles and binary centres of mass, which is
    what kira places in GRAPE j-memory) and the binaries (index of the
    top-level node, component masses, relative orbit)."""
    _, x, v, model = makeking(n, w0, seed)
    m = power_law_masses(n, seed + 1, 1.0, mass_ratio, exponent)
    prim, msec = add_secondaries(m, bin_frac, seed + 2)
    m1 = m[prim].copy()
    mtop = m.copy()
    mtop[prim] += msec
    mt0 = float(mtop.sum())
    mtop, x, v = scale_nbody(mtop, x, v, potential_fn)
    m1 = m1 * (1.0 / mt0)
    m2 = msec * (1.0 / mt0)
    a, e, dx, dv = binary_orbits(m1, m2, seed + 3, a_range[0], a_range[1])
    return dict(m=mtop, x=x, v=v, model=model, bin_index=prim, m1=m1, m2=m2, a=a, e=e,
                dx=dx, dv=dv)
