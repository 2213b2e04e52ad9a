/*
 * g6_oracle.c -- CPU oracle for the GRAPE-6 force path (TEST INFRASTRUCTURE).
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load this code, and only to check results.  The product library
 * never links or calls it.
 *
 * Build: see oracle/Makefile (always -O2 -ffp-contract=off: every FP64
 * expression rounds as written, in the order written).
 *
 * Parity pins (tests/test_oracle.py): two-body closed form, Newton's third
 * law, Galilean/translation invariance, jerk = d(acc)/dt by finite
 * differences, long-double brute force, the IEEE GRAPE-4 simulator built
 * from the reference sources (oracle/_ref), Taylor-series identities of the
 * predictor, brute-force neighbour sets, ran2 against an independent
 * big-integer restatement of its two LCGs and shuffle, and Plummer
 * invariants.  The shared-step integrator (or_hermite_shared) is pinned in
 * tests/test_ens_cpu.py: a Kepler orbit closes after one period, the error
 * falls as eta^4, momentum is conserved.  The block-step integrator
 * (or_hermite_block) is pinned there too: bit for bit against kira's
 * block-step rules restated independently in vectorised numpy
 * (starlab_amd/blockstep.py on an FP64 host backend), bit for bit against
 * the shared-step integrator where every particle keeps the same step (two
 * equal masses; a step cap below every natural step), a Kepler orbit closes,
 * and the error falls as eta^4.
 */
#include "g6_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* ran2: Numerical Recipes (2nd ed.) long-period generator, as used by      */
/* Starlab's randinter (src/std/randinter.C:257-324).  Constants are the    */
/* published ones; EPS is Starlab's 1e-14 (randinter.C:255), not NR's.       */
/* ------------------------------------------------------------------------ */

#define R2_M1   2147483563
#define R2_M2   2147483399
#define R2_A1   40014
#define R2_A2   40692
#define R2_Q1   53668
#define R2_Q2   52774
#define R2_R1   12211
#define R2_R2   3791
#define R2_TAB  32
#define R2_DIV  (1 + (R2_M1 - 1) / R2_TAB)
#define R2_MAX  (1.0 - 1.e-14)

static int r2_state = 0;          /* randx in randinter.C:49 */
static int r2_state2 = 123456789; /* idum2 */
static int r2_y = 0;
static int r2_v[R2_TAB];
static int r2_calls = 0;

static int lcg_step(int s, int a, int q, int r, int m)
{
    /* Schrage's method: a*s mod m without overflow */
    int k = s / q;
    s = a * (s - k * q) - k * r;
    if (s < 0) s += m;
    return s;
}

static double ran2_next(void)
{
    if (r2_state <= 0) {                 /* (re)initialise the shuffle table */
        int s = -r2_state;
        if (s < 1) s = 1;
        r2_state2 = s;
        for (int t = R2_TAB + 7; t >= 0; t--) {
            s = lcg_step(s, R2_A1, R2_Q1, R2_R1, R2_M1);
            if (t < R2_TAB) r2_v[t] = s;
        }
        r2_y = r2_v[0];
        r2_state = s;
    }
    r2_state  = lcg_step(r2_state,  R2_A1, R2_Q1, R2_R1, R2_M1);
    r2_state2 = lcg_step(r2_state2, R2_A2, R2_Q2, R2_R2, R2_M2);
    int slot = r2_y / R2_DIV;
    r2_y = r2_v[slot] - r2_state2;
    r2_v[slot] = r2_state;
    if (r2_y < 1) r2_y += R2_M1 - 1;
    double u = (1.0 / R2_M1) * r2_y;
    return u > R2_MAX ? R2_MAX : u;
}

int or_ran2_seed(int seed)
{
    /* srandinter(seed, 0): randx = -|seed| forces re-initialisation */
    int s = abs(seed);
    r2_state = -s;
    r2_calls = 0;
    return s;
}

double or_randinter(double a, double b)
{
    r2_calls++;
    return a + (b - a) * ran2_next();
}

int or_n_rand(void) { return r2_calls; }

int or_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Plummer model (makeplummer.C:55-214).                                     */
/* ------------------------------------------------------------------------ */

static void plummer_energies(int n, const double *m, const double *x,
                             const double *v, double *epot, double *ekin)
{
    or_energies(n, m, x, v, 0.0, epot, ekin);
}

void or_makeplummer(int n, int seed, int scale, double *m, double *x, double *v)
{
    const double PI = 3.14159265358979323846;
    const double TWO_PI = 2.0 * PI;
    double mfrac = 0.999;                      /* MFRAC_DEFAULT, :50 */
    double partmass = 1.0 / (double)n;

    or_ran2_seed(seed);

    for (int i = 0; i < n; i++) {
        m[i] = partmass;
        double m_min = (i * mfrac) / n;
        double m_max = ((i + 1) * mfrac) / n;
        double rrrr = or_randinter(m_min, m_max);
        double radius = 1 / sqrt(pow(rrrr, -2.0 / 3.0) - 1);
        double theta = acos(or_randinter(-1.0, 1.0));
        double phi = or_randinter(0.0, TWO_PI);
        x[3*i+0] = radius * sin(theta) * cos(phi);
        x[3*i+1] = radius * sin(theta) * sin(phi);
        x[3*i+2] = radius * cos(theta);

        double xx = 0.0, yy = 0.1;             /* von Neumann rejection */
        while (yy > xx * xx * pow(1.0 - xx * xx, 3.5)) {
            xx = or_randinter(0.0, 1.0);
            yy = or_randinter(0.0, 0.1);
        }
        double velocity = xx * sqrt(2.0) * pow(1.0 + radius * radius, -0.25);
        theta = acos(or_randinter(-1.0, 1.0));
        phi = or_randinter(0.0, TWO_PI);
        v[3*i+0] = velocity * sin(theta) * cos(phi);
        v[3*i+1] = velocity * sin(theta) * sin(phi);
        v[3*i+2] = velocity * cos(theta);
    }

    double xfac = 1 / 1.695;                   /* :186-193 */
    double vfac = 1 / sqrt(xfac);
    for (int i = 0; i < 3 * n; i++) { x[i] *= xfac; v[i] *= vfac; }

    /* to_com(): move to the centre-of-mass frame */
    double mt = 0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        mt += m[i];
        for (int k = 0; k < 3; k++) {
            cx[k] += m[i] * x[3*i+k];
            cv[k] += m[i] * v[3*i+k];
        }
    }
    for (int k = 0; k < 3; k++) { cx[k] /= mt; cv[k] /= mt; }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) { x[3*i+k] -= cx[k]; v[3*i+k] -= cv[k]; }

    if (scale && n > 1) {
        /* scale_virial(-0.5) then scale_energy(-0.25), eps = 0 (scale.C:157-190) */
        double pot, kin;
        plummer_energies(n, m, x, v, &pot, &kin);
        double q = -0.5;
        double vscale = sqrt(q * pot / kin);
        for (int i = 0; i < 3 * n; i++) v[i] *= vscale;
        kin = q * pot;
        double energy = kin + pot;
        if (energy < 0) {
            double xscale = energy / -0.25;
            double vs2 = sqrt(1. / xscale);
            for (int i = 0; i < 3 * n; i++) { x[i] *= xscale; v[i] *= vs2; }
        }
    }

    /* reshuffle_all (:236-253): swap each slot with a random one */
    for (int i = 0; i < n; i++) {
        int j = (int)or_randinter(0, n);
        if (j != i) {
            double t = m[i]; m[i] = m[j]; m[j] = t;
            for (int k = 0; k < 3; k++) {
                t = x[3*i+k]; x[3*i+k] = x[3*j+k]; x[3*j+k] = t;
                t = v[3*i+k]; v[3*i+k] = v[3*j+k]; v[3*j+k] = t;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* j predictor.  Fields follow g6_set_j_particle_: a2 = a/2, j6 = j/6,       */
/* k18 = a''/18 (hdyn_grape6.C:246-249, hdyn_ev.C:1076).  With s = ti - tj,  */
/* the 4th-order form is the reference GRAPE simulator's Horner scheme       */
/* (src/harpsims/harp3fastsim.c:251-265):                                    */
/*   xp = ((j6*s + a2)*s + v)*s + x                                          */
/*   vp = (j6*(1.5 s) + a2)*(2 s) + v                                        */
/* 5th order replaces j6 by (k18*(0.75 s) + j6) in xp and by (k18*s + j6)    */
/* in vp, i.e. the Taylor terms (3/4) k18 s^4 and 3 k18 s^3; with k18 = 0    */
/* both orders round identically.                                            */
/* ------------------------------------------------------------------------ */

void or_predict(int nj, const double *x, const double *v, const double *a2,
                const double *j6, const double *k18, const double *tj,
                double ti, int order, double *xp, double *vp)
{
    for (int j = 0; j < nj; j++) {
        double s = ti - tj[j];
        double s1 = 1.5 * s, s2 = 2.0 * s;
        for (int k = 0; k < 3; k++) {
            int q = 3 * j + k;
            double cx = j6[q], cv = j6[q];
            if (order >= 5) {
                cx = k18[q] * (0.75 * s) + j6[q];
                cv = k18[q] * s + j6[q];
            }
            xp[q] = ((cx * s + a2[q]) * s + v[q]) * s + x[q];
            vp[q] = (cv * s1 + a2[q]) * s2 + v[q];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Direct sum: accumulate_acc_and_jerk (hdyn_ev.C:2019-2036), evaluated in   */
/* the same order; d = x_j - x_i; dot products left to right                 */
/* (starlab_vector.h:68-71).                                                 */
/* ------------------------------------------------------------------------ */

void or_force(int ni, const double *xi, const double *vi, const int *excl,
              int nj, const double *xj, const double *vj, const double *mj,
              const int *jindex, double eps2,
              double *acc, double *jerk, double *pot, int *nn, double *dnn2)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; i++) {
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, p = 0;
        double best = INFINITY;
        int besti = -1;
        const double x0 = xi[3*i], y0 = xi[3*i+1], z0 = xi[3*i+2];
        const double u0 = vi[3*i], v0 = vi[3*i+1], w0 = vi[3*i+2];
        for (int j = 0; j < nj; j++) {
            if (jindex[j] == excl[i]) continue;
            double dx = xj[3*j] - x0, dy = xj[3*j+1] - y0, dz = xj[3*j+2] - z0;
            double du = vj[3*j] - u0, dv = vj[3*j+1] - v0, dw = vj[3*j+2] - w0;
            double d2 = dx * dx + dy * dy + dz * dz;
            double r2inv = 1.0 / (d2 + eps2);
            double a3 = -3.0 * (dx * du + dy * dv + dz * dw) * r2inv;
            double mrinv = mj[j] * sqrt(r2inv);
            p -= mrinv;
            double mr3inv = mrinv * r2inv;
            ax += mr3inv * dx; ay += mr3inv * dy; az += mr3inv * dz;
            jx += mr3inv * (du + a3 * dx);
            jy += mr3inv * (dv + a3 * dy);
            jz += mr3inv * (dw + a3 * dz);
            if (d2 < best) { best = d2; besti = jindex[j]; }   /* strict <, hdyn_inline.C:44 */
        }
        acc[3*i] = ax; acc[3*i+1] = ay; acc[3*i+2] = az;
        jerk[3*i] = jx; jerk[3*i+1] = jy; jerk[3*i+2] = jz;
        pot[i] = p;
        if (nn) nn[i] = besti;
        if (dnn2) dnn2[i] = best;
    }
}

void or_force_ld(const double *xi, const double *vi, int excl,
                 int nj, const double *xj, const double *vj, const double *mj,
                 const int *jindex, double eps2,
                 double *acc, double *jerk, double *pot)
{
    long double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, p = 0;
    for (int j = 0; j < nj; j++) {
        if (jindex[j] == excl) continue;
        long double d[3], w[3];
        for (int k = 0; k < 3; k++) {
            d[k] = (long double)xj[3*j+k] - (long double)xi[k];
            w[k] = (long double)vj[3*j+k] - (long double)vi[k];
        }
        long double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2] + (long double)eps2;
        long double rv = d[0]*w[0] + d[1]*w[1] + d[2]*w[2];
        long double rinv = 1.0L / sqrtl(r2);
        long double mr = (long double)mj[j] * rinv;
        long double mr3 = mr * rinv * rinv;
        p -= mr;
        for (int k = 0; k < 3; k++) {
            a[k] += mr3 * d[k];
            jk[k] += mr3 * (w[k] - 3.0L * rv / r2 * d[k]);
        }
    }
    for (int k = 0; k < 3; k++) { acc[k] = (double)a[k]; jerk[k] = (double)jk[k]; }
    *pot = (double)p;
}

/* ------------------------------------------------------------------------ */
/* Neighbour spheres.                                                        */
/* ------------------------------------------------------------------------ */

void or_neighbours(int ni, const double *xi, const double *h2, const int *excl,
                   int include_self, int nj, const double *xj, const int *jindex,
                   int cap, int *counts, int *lists)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; i++) {
        int c = 0;
        if (h2[i] > 0) {
            for (int j = 0; j < nj; j++) {
                if (!include_self && jindex[j] == excl[i]) continue;
                double dx = xj[3*j] - xi[3*i];
                double dy = xj[3*j+1] - xi[3*i+1];
                double dz = xj[3*j+2] - xi[3*i+2];
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < h2[i]) {
                    if (c < cap) lists[(long)i * cap + c] = jindex[j];
                    c++;
                }
            }
        }
        counts[i] = c;
    }
}

/* ------------------------------------------------------------------------ */
/* GRAPE-4 semantics: the pair loop of the reference's IEEE GRAPE-4          */
/* simulator, src/harpsims/harp3fastsim.c:268-322 (per-i softening and       */
/* neighbour radius doubled as at :240-241; a pair with r^2 = 0 is skipped,  */
/* :304), in its operation order, plus pot = -sum m/r over the same pairs    */
/* (the simulator leaves pot at 0).  Lists hold j addresses in ascending     */
/* order, the i itself included when r^2 = 2 eps2 < 2 h2.                    */
/* ------------------------------------------------------------------------ */

void or_g4_force(int ni, const double *xi, const double *vi, const double *eps2,
                 const double *h2, int nj, const double *xj, const double *vj,
                 const double *mj, double *acc, double *jerk, double *pot,
                 int cap, int *counts, int *lists)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; i++) {
        const double e = 2.0 * eps2[i], h = 2.0 * h2[i];
        const double x0 = xi[3*i], y0 = xi[3*i+1], z0 = xi[3*i+2];
        const double u0 = vi[3*i], v0 = vi[3*i+1], w0 = vi[3*i+2];
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, p = 0;
        int c = 0;
        for (int j = 0; j < nj; j++) {
            double dx = xj[3*j] - x0, dy = xj[3*j+1] - y0, dz = xj[3*j+2] - z0;
            double du = vj[3*j] - u0, dv = vj[3*j+1] - v0, dw = vj[3*j+2] - w0;
            double r2 = e + dx * dx + dy * dy + dz * dz;
            double rv = dx * du + dy * dv + dz * dw;
            if (r2 < h) {
                if (counts && c < cap) lists[(long)i * cap + c] = j;
                c++;
            }
            if (r2 > 0.0) {
                double r2i = 1.0 / r2;
                double rinv = sqrt(r2i);
                double r3i = mj[j] * r2i * rinv;
                double q = 3.0 * rv * r2i;
                ax += dx * r3i; ay += dy * r3i; az += dz * r3i;
                jx += (du - dx * q) * r3i;
                jy += (dv - dy * q) * r3i;
                jz += (dw - dz * q) * r3i;
                p -= mj[j] * rinv;
            }
        }
        acc[3*i] = ax; acc[3*i+1] = ay; acc[3*i+2] = az;
        jerk[3*i] = jx; jerk[3*i+1] = jy; jerk[3*i+2] = jz;
        pot[i] = p;
        if (counts) counts[i] = c;
    }
}

/* ------------------------------------------------------------------------ */
/* Local density from the k-th nearest neighbour: compute_density,           */
/* src/node/dyn/util/compute_density.C:69-226, whole-system mode (every      */
/* other particle in ascending order, :155-189).  d[0] = 0 stands for the    */
/* particle itself; a new distance goes after equal ones; density = masses   */
/* of neighbours 1..k-1 over (4/3) pi d_k^3.                                 */
/* ------------------------------------------------------------------------ */

#define OR_VERY_LARGE 1.0e300      /* VERY_LARGE_NUMBER, include/stdinc.h:206 */
#define OR_PI 3.14159265358979323846

void or_density(int n, const double *m, const double *x, int i_first, int ni, int k,
                double *density, double *dk2)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int ii = 0; ii < ni; ii++) {
        const int i = i_first + ii;
        double nd[64], nm[64];
        nd[0] = 0.0; nm[0] = m[i];
        for (int q = 1; q <= k; q++) { nd[q] = OR_VERY_LARGE; nm[q] = 0.0; }
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 < nd[k])) continue;
            for (int q = k - 1; q >= 0; q--) {
                if (d2 >= nd[q]) {
                    for (int r = k; r > q + 1; r--) { nd[r] = nd[r-1]; nm[r] = nm[r-1]; }
                    nd[q+1] = d2; nm[q+1] = m[j];
                    break;
                }
            }
        }
        double rho = 0.0;
        int kk = k;
        if (nd[kk] > 0) {
            while (kk > 0 && nd[kk] >= OR_VERY_LARGE) kk--;
            if (kk > 1 && nd[kk] > 0 && nd[kk] < OR_VERY_LARGE) {
                double volume = (4.0 / 3.0) * OR_PI * pow(nd[kk], 1.5);
                if (volume > 0) {
                    double mass = 0;
                    for (int q = 1; q < kk; q++) mass += nm[q];
                    rho = mass / volume;
                }
            }
        }
        density[ii] = rho;
        dk2[ii] = nd[k];
    }
}

/* ------------------------------------------------------------------------ */
/* Energies (each pair once).                                                */
/* ------------------------------------------------------------------------ */

void or_energies(int n, const double *m, const double *x, const double *v,
                 double eps2, double *epot, double *ekin)
{
    double pot = 0, kin = 0;
#pragma omp parallel for reduction(+:pot) schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        double pi = 0;
        for (int j = i + 1; j < n; j++) {
            double dx = x[3*j] - x[3*i], dy = x[3*j+1] - x[3*i+1], dz = x[3*j+2] - x[3*i+2];
            pi -= m[j] / sqrt(dx * dx + dy * dy + dz * dz + eps2);
        }
        pot += m[i] * pi;
    }
    for (int i = 0; i < n; i++)
        kin += m[i] * (v[3*i] * v[3*i] + v[3*i+1] * v[3*i+1] + v[3*i+2] * v[3*i+2]);
    *epot = pot;
    *ekin = 0.5 * kin;
}

/* ------------------------------------------------------------------------ */
/* Shared-step Hermite integrator (checker of the ensemble engine).          */
/* ------------------------------------------------------------------------ */

static void hs_force(int n, const double *m, const double *x, const double *v, double eps2,
                     double *a, double *j, double *pot)
{
    for (int i = 0; i < n; i++) {
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, p = 0;
        for (int k = 0; k < n; k++) {
            if (k == i) continue;
            double dx = x[3*k] - x[3*i], dy = x[3*k+1] - x[3*i+1], dz = x[3*k+2] - x[3*i+2];
            double du = v[3*k] - v[3*i], dv = v[3*k+1] - v[3*i+1], dw = v[3*k+2] - v[3*i+2];
            double d2 = dx * dx + dy * dy + dz * dz;
            double r2inv = 1.0 / (d2 + eps2);
            double a3 = -3.0 * (dx * du + dy * dv + dz * dw) * r2inv;
            double mrinv = m[k] * sqrt(r2inv);
            p -= mrinv;
            double mr3inv = mrinv * r2inv;
            ax += mr3inv * dx; ay += mr3inv * dy; az += mr3inv * dz;
            jx += mr3inv * (du + a3 * dx);
            jy += mr3inv * (dv + a3 * dy);
            jz += mr3inv * (dw + a3 * dz);
        }
        a[3*i] = ax; a[3*i+1] = ay; a[3*i+2] = az;
        j[3*i] = jx; j[3*i+1] = jy; j[3*i+2] = jz;
        pot[i] = p;
    }
}

static double hs_energy(int n, const double *m, const double *v, const double *pot)
{
    double kin = 0, pe = 0;
    for (int i = 0; i < n; i++) {
        kin += m[i] * (v[3*i] * v[3*i] + v[3*i+1] * v[3*i+1] + v[3*i+2] * v[3*i+2]);
        pe += m[i] * pot[i];
    }
    return 0.5 * kin + 0.5 * pe;
}

static double hs_pow2_below(double x)
{
    int e;
    frexp(x, &e);
    return ldexp(1.0, e - 1);
}

int or_hermite_shared(int n, const double *m, double *x, double *v, double t_end,
                      double eta, double eps2, double dt_max, long long max_steps,
                      double *t_final, long long *steps, double *e0, double *e1)
{
    const double ONE3 = 1.0 / 3.0, ONE12 = 1.0 / 12.0;
    double *a = calloc(3 * (size_t)n, sizeof(double)), *jk = calloc(3 * (size_t)n, sizeof(double));
    double *a1 = calloc(3 * (size_t)n, sizeof(double)), *j1 = calloc(3 * (size_t)n, sizeof(double));
    double *xp = calloc(3 * (size_t)n, sizeof(double)), *vp = calloc(3 * (size_t)n, sizeof(double));
    double *pot = calloc((size_t)n, sizeof(double));
    int rc = 0;
    hs_force(n, m, x, v, eps2, a, jk, pot);
    *e0 = hs_energy(n, m, v, pot);
    /* first step (set_first_timestep), minimum over particles */
    double nat = INFINITY;
    for (int i = 0; i < n; i++) {
        double eta_init = eta / 16;
        double asq = a[3*i]*a[3*i] + a[3*i+1]*a[3*i+1] + a[3*i+2]*a[3*i+2];
        double jsq = jk[3*i]*jk[3*i] + jk[3*i+1]*jk[3*i+1] + jk[3*i+2]*jk[3*i+2];
        double dt_adot = 1e300, dt_ff = 1e300;
        if (jsq > 0) dt_adot = asq / jsq;
        if (pot[i] < 0 && asq > 0) dt_ff = -pot[i] / asq;
        double c = eta_init * sqrt(fmin(dt_adot, dt_ff));
        c = fmin(c, eta_init);
        if (c < nat) nat = c;
    }
    double dt_cap = fmin(dt_max, t_end);
    double t = 0, dt = hs_pow2_below(fmin(nat, dt_cap));
    if (t + dt > t_end) dt = t_end - t;
    long long ns = 0;
    while (t < t_end) {
        for (int q = 0; q < 3 * n; q++) {       /* predict_loworder */
            double dt3 = dt * ONE3, dt2 = dt * 0.5;
            xp[q] = ((jk[q] * dt3 + a[q]) * dt2 + v[q]) * dt + x[q];
            vp[q] = (jk[q] * dt2 + a[q]) * dt + v[q];
        }
        hs_force(n, m, xp, vp, eps2, a1, j1, pot);
        nat = INFINITY;
        for (int i = 0; i < n; i++) {
            double bt2[3], at3[3];
            for (int k = 0; k < 3; k++) {
                int q = 3 * i + k;
                bt2[k] = -3 * (a[q] - a1[q]) - dt * (2 * jk[q] + j1[q]);
                at3[k] = 2 * (a[q] - a1[q]) + dt * (jk[q] + j1[q]);
            }
            for (int k = 0; k < 3; k++) {
                int q = 3 * i + k;
                x[q] = xp[q] + (0.05 * at3[k] + ONE12 * bt2[k]) * dt * dt;
                v[q] = vp[q] + (0.25 * at3[k] + ONE3 * bt2[k]) * dt;
                a[q] = a1[q];
                jk[q] = j1[q];
            }
            double dt2 = dt * dt, dt4 = dt2 * dt2, dt6 = dt4 * dt2;
            double l2 = 36 * (at3[0]*at3[0] + at3[1]*at3[1] + at3[2]*at3[2]) / dt6;
            double k2 = 4 * (bt2[0]*bt2[0] + bt2[1]*bt2[1] + bt2[2]*bt2[2]) / dt4;
            double j2 = j1[3*i]*j1[3*i] + j1[3*i+1]*j1[3*i+1] + j1[3*i+2]*j1[3*i+2];
            double asq = a1[3*i]*a1[3*i] + a1[3*i+1]*a1[3*i+1] + a1[3*i+2]*a1[3*i+2];
            double c = eta * sqrt((sqrt(asq * k2) + j2) / (sqrt(j2 * l2) + k2));
            if (!(c > 0)) c = INFINITY;
            if (c < nat) nat = c;
        }
        t += dt;
        ns++;
        if (t >= t_end) break;
        double T = dt, nd = T;
        if (nat < T) nd = 0.5 * T;
        else if (nat >= 2 * T && fmod(t, 2 * T) == 0.0 && 2 * T <= dt_cap) nd = 2 * T;
        if (t + nd > t_end) nd = t_end - t;
        dt = nd;
        if (!(dt > 1e-30) || ns > max_steps) { rc = 1; break; }
    }
    hs_force(n, m, x, v, eps2, a1, j1, pot);
    *e1 = hs_energy(n, m, v, pot);
    *t_final = t;
    *steps = ns;
    free(a); free(jk); free(a1); free(j1); free(xp); free(vp); free(pot);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Block-step Hermite integrator (checker of the ensemble engine's block     */
/* mode): kira's individual power-of-two steps.                              */
/* ------------------------------------------------------------------------ */

/* One system from t = 0 to t_end.  Each block is every particle with the
 * smallest t + dt (fast_get_nodes_to_move, hdyn_schedule.C:641-745); all
 * particles are predicted to the block time with predict_loworder
 * (_dyn_.h:290-295), the block's forces are summed over every other particle
 * in index order with the pair law of hs_force, and each block particle is
 * corrected (hdyn_ev.C:1235-1250, :1717-1718) and given its own next step:
 * the Aarseth criterion (:627-655), halved once if it falls short, doubled
 * when the block time is commensurate with the doubled step (:866-887),
 * never above min(dt_max, t_end).  First steps: set_first_timestep
 * (:377-410) per particle.  A step that would pass t_end is cut to end
 * there. */
int or_hermite_block(int n, const double *m, double *x, double *v, double t_end,
                     double eta, double eps2, double dt_max, long long max_blocks,
                     double *t_final, long long *blocks, long long *psteps,
                     double *e0, double *e1)
{
    const double ONE3 = 1.0 / 3.0, ONE12 = 1.0 / 12.0;
    double *a = calloc(3 * (size_t)n, sizeof(double)), *jk = calloc(3 * (size_t)n, sizeof(double));
    double *a1 = calloc(3 * (size_t)n, sizeof(double)), *j1 = calloc(3 * (size_t)n, sizeof(double));
    double *xp = calloc(3 * (size_t)n, sizeof(double)), *vp = calloc(3 * (size_t)n, sizeof(double));
    double *pot = calloc((size_t)n, sizeof(double)), *p1 = calloc((size_t)n, sizeof(double));
    double *t = calloc((size_t)n, sizeof(double)), *dt = calloc((size_t)n, sizeof(double));
    int *act = calloc((size_t)n, sizeof(int));
    int rc = 0;
    hs_force(n, m, x, v, eps2, a, jk, pot);
    *e0 = hs_energy(n, m, v, pot);
    const double dt_cap = fmin(dt_max, t_end);
    for (int i = 0; i < n; i++) {
        double eta_init = eta / 16;
        double asq = a[3*i]*a[3*i] + a[3*i+1]*a[3*i+1] + a[3*i+2]*a[3*i+2];
        double jsq = jk[3*i]*jk[3*i] + jk[3*i+1]*jk[3*i+1] + jk[3*i+2]*jk[3*i+2];
        double dt_adot = 1e300, dt_ff = 1e300;
        if (jsq > 0) dt_adot = asq / jsq;
        if (pot[i] < 0 && asq > 0) dt_ff = -pot[i] / asq;
        double c = eta_init * sqrt(fmin(dt_adot, dt_ff));
        c = fmin(c, eta_init);
        dt[i] = hs_pow2_below(fmin(c, dt_cap));
        if (!(dt[i] > 1e-30)) rc = 1;
    }
    long long nb = 0, ps = 0;
    double tlast = 0;
    while (!rc) {
        double tn = INFINITY;
        for (int i = 0; i < n; i++)
            if (t[i] < t_end && t[i] + dt[i] < tn) tn = t[i] + dt[i];
        if (tn == INFINITY) break;                 /* every particle is at t_end */
        int na = 0;
        for (int i = 0; i < n; i++)
            if (t[i] < t_end && t[i] + dt[i] == tn) act[na++] = i;
        for (int i = 0; i < n; i++) {             /* predict_loworder to the block time */
            double d = tn - t[i], d3 = d * ONE3, d2 = d * 0.5;
            for (int k = 0; k < 3; k++) {
                int q = 3 * i + k;
                xp[q] = ((jk[q] * d3 + a[q]) * d2 + v[q]) * d + x[q];
                vp[q] = (jk[q] * d2 + a[q]) * d + v[q];
            }
        }
        for (int b = 0; b < na; b++) {            /* forces on the block */
            int i = act[b];
            double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, p = 0;
            for (int k = 0; k < n; k++) {
                if (k == i) continue;
                double dx = xp[3*k] - xp[3*i], dy = xp[3*k+1] - xp[3*i+1], dz = xp[3*k+2] - xp[3*i+2];
                double du = vp[3*k] - vp[3*i], dv = vp[3*k+1] - vp[3*i+1], dw = vp[3*k+2] - vp[3*i+2];
                double d2 = dx * dx + dy * dy + dz * dz;
                double r2inv = 1.0 / (d2 + eps2);
                double a3 = -3.0 * (dx * du + dy * dv + dz * dw) * r2inv;
                double mrinv = m[k] * sqrt(r2inv);
                p -= mrinv;
                double mr3inv = mrinv * r2inv;
                ax += mr3inv * dx; ay += mr3inv * dy; az += mr3inv * dz;
                jx += mr3inv * (du + a3 * dx);
                jy += mr3inv * (dv + a3 * dy);
                jz += mr3inv * (dw + a3 * dz);
            }
            a1[3*b] = ax; a1[3*b+1] = ay; a1[3*b+2] = az;
            j1[3*b] = jx; j1[3*b+1] = jy; j1[3*b+2] = jz;
            p1[b] = p;
        }
        for (int b = 0; b < na; b++) {            /* correct, next step */
            int i = act[b];
            double d = dt[i];
            double bt2[3], at3[3];
            for (int k = 0; k < 3; k++) {
                int q = 3 * i + k, r = 3 * b + k;
                bt2[k] = -3 * (a[q] - a1[r]) - d * (2 * jk[q] + j1[r]);
                at3[k] = 2 * (a[q] - a1[r]) + d * (jk[q] + j1[r]);
            }
            for (int k = 0; k < 3; k++) {
                int q = 3 * i + k, r = 3 * b + k;
                x[q] = xp[q] + (0.05 * at3[k] + ONE12 * bt2[k]) * d * d;
                v[q] = vp[q] + (0.25 * at3[k] + ONE3 * bt2[k]) * d;
                a[q] = a1[r];
                jk[q] = j1[r];
            }
            pot[i] = p1[b];
            double d2 = d * d, d4 = d2 * d2, d6 = d4 * d2;
            double l2 = 36 * (at3[0]*at3[0] + at3[1]*at3[1] + at3[2]*at3[2]) / d6;
            double k2 = 4 * (bt2[0]*bt2[0] + bt2[1]*bt2[1] + bt2[2]*bt2[2]) / d4;
            const double *jn = &j1[3*b], *an = &a1[3*b];
            double j2 = jn[0]*jn[0] + jn[1]*jn[1] + jn[2]*jn[2];
            double asq = an[0]*an[0] + an[1]*an[1] + an[2]*an[2];
            double nat = eta * sqrt((sqrt(asq * k2) + j2) / (sqrt(j2 * l2) + k2));
            if (!(nat > 0)) nat = INFINITY;
            double T = d, nd = T;
            if (nat < T) nd = 0.5 * T;
            else if (nat >= 2 * T && fmod(tn, 2 * T) == 0.0 && 2 * T <= dt_cap) nd = 2 * T;
            if (tn + nd > t_end) nd = t_end - tn;
            t[i] = tn;
            dt[i] = nd;
            if (tn < t_end && !(nd > 1e-30)) rc = 1;
        }
        tlast = tn;
        nb++;
        ps += na;
        if (nb > max_blocks) rc = 1;
    }
    hs_force(n, m, x, v, eps2, a1, j1, pot);
    *e1 = hs_energy(n, m, v, pot);
    *t_final = tlast;
    *blocks = nb;
    *psteps = ps;
    free(a); free(jk); free(a1); free(j1); free(xp); free(vp); free(pot); free(p1);
    free(t); free(dt); free(act);
    return rc;
}
