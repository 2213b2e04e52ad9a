/*
 * abi_bench.c -- per-block cost of the GRAPE-6 ABI as kira drives it, from a
 * compiled C caller (no Python in the loop).
 *
 *     abi_bench NJ BLOCKS NI [NI ...]
 *
 * Loads NJ j-particles (a uniform sphere, equal masses, random velocities),
 * then for each block size NI runs BLOCKS blocks of kira's per-block
 * sequence (hdyn_grape6.C:3067-3338):
 *
 *   1. g6_set_j_particle_ for each particle of the previous block (kira
 *      re-sends the nodes it just advanced, :3112-3113), with a new tj;
 *   2. g6_set_ti_;
 *   3. g6calc_firsthalf_ / g6calc_lasthalf2_ on NI i-particles (NI <=
 *      g6_npipes_()), positions predicted on the host as kira does.
 *
 * With ABI_BENCH_H2 set in the environment, every i gets that neighbour
 * radius squared and each block also calls g6_read_neighbour_list_ and one
 * g6_get_neighbour_list_ per i, as kira does for nodes with h2 > 0
 * (hdyn_grape6.C:2062-2066, 2302-2306); their time is reported separately.
 *
 * Prints one JSON line per NI: median and mean microseconds per block and per
 * phase, and pair interactions per second.  Exit status 0 on success.
 *
 * With ABI_BENCH_DUMP naming a file, the last block of each NI is appended
 * to it (native byte order) so that a checker can recompute it:
 *   int32 nj, ni; float64 ti, eps2, m;  float64 tj[nj], x[nj][3], v[nj][3];
 *   int32 index[ni]; float64 xi[ni][3], vi[ni][3];
 *   float64 acc[ni][3], jerk[ni][3], pot[ni]; int32 nn[ni].
 * The j-particles carry zero a2, j6 and k18, so their predicted state at ti
 * is x + (ti - tj) v, v.
 */
#define _POSIX_C_SOURCE 199309L
#include "grape6.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_us(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e6 + t.tv_nsec * 1e-3;
}

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int dump_block(const char *path, int nj, int ni, double ti, double eps2, double m,
                      const double *tj, const double (*x)[3], const double (*v)[3],
                      const int *index, const double (*xi)[3], const double (*vi)[3],
                      const double (*acc)[3], const double (*jerk)[3], const double *pot,
                      const int *nn)
{
    FILE *f = fopen(path, "ab");
    if (!f) return 1;
    const int head[2] = {nj, ni};
    const double par[3] = {ti, eps2, m};
    size_t ok = fwrite(head, sizeof(int), 2, f) == 2;
    ok &= fwrite(par, sizeof(double), 3, f) == 3;
    ok &= fwrite(tj, sizeof(double), nj, f) == (size_t)nj;
    ok &= fwrite(x, sizeof(double[3]), nj, f) == (size_t)nj;
    ok &= fwrite(v, sizeof(double[3]), nj, f) == (size_t)nj;
    ok &= fwrite(index, sizeof(int), ni, f) == (size_t)ni;
    ok &= fwrite(xi, sizeof(double[3]), ni, f) == (size_t)ni;
    ok &= fwrite(vi, sizeof(double[3]), ni, f) == (size_t)ni;
    ok &= fwrite(acc, sizeof(double[3]), ni, f) == (size_t)ni;
    ok &= fwrite(jerk, sizeof(double[3]), ni, f) == (size_t)ni;
    ok &= fwrite(pot, sizeof(double), ni, f) == (size_t)ni;
    ok &= fwrite(nn, sizeof(int), ni, f) == (size_t)ni;
    return (fclose(f) != 0 || !ok) ? 1 : 0;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static double median(double *v, int n)
{
    qsort(v, n, sizeof(double), cmp_double);
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s NJ BLOCKS NI [NI ...]\n", argv[0]);
        return 2;
    }
    const int nj = atoi(argv[1]), nblocks = atoi(argv[2]);
    if (nj < 2 || nblocks < 1) return 2;
    int cl = 0;
    const int npipes = g6_npipes_();
    if (g6_open_(&cl) != 0) {
        fprintf(stderr, "g6_open_ failed\n");
        return 1;
    }
    double (*x)[3] = malloc(sizeof(double[3]) * nj);
    double (*v)[3] = malloc(sizeof(double[3]) * nj);
    double *tj = malloc(sizeof(double) * nj);
    const double m = 1.0 / nj, dtj = 1.0 / 1024;
    double zero[3] = {0, 0, 0};
    for (int k = 0; k < nj; ++k) {
        double r2;
        do {
            for (int c = 0; c < 3; ++c) x[k][c] = 2 * urand() - 1;
            r2 = x[k][0] * x[k][0] + x[k][1] * x[k][1] + x[k][2] * x[k][2];
        } while (r2 > 1 || r2 == 0);
        for (int c = 0; c < 3; ++c) v[k][c] = 0.5 * (2 * urand() - 1);
        tj[k] = 0;
    }
    double t0 = now_us();
    for (int k = 0; k < nj; ++k) {
        double mm = m, t = tj[k], d = dtj;
        int a = k, idx = k;
        if (g6_set_j_particle_(&cl, &a, &idx, &t, &d, &mm, zero, zero, zero, v[k], x[k])) {
            fprintf(stderr, "g6_set_j_particle_ failed\n");
            return 1;
        }
    }
    const double load_us = now_us() - t0;

    int *index = malloc(sizeof(int) * npipes);
    double (*xi)[3] = malloc(sizeof(double[3]) * npipes);
    double (*vi)[3] = malloc(sizeof(double[3]) * npipes);
    double (*acc)[3] = malloc(sizeof(double[3]) * npipes);
    double (*jerk)[3] = malloc(sizeof(double[3]) * npipes);
    double *pot = malloc(sizeof(double) * npipes);
    double *h2 = calloc(npipes, sizeof(double));
    int *nn = malloc(sizeof(int) * npipes);
    const char *h2env = getenv("ABI_BENCH_H2");
    const double h2val = h2env ? atof(h2env) : 0.0;
    const char *dump = getenv("ABI_BENCH_DUMP");
    for (int k = 0; k < npipes; ++k) h2[k] = h2val;
    int maxlen = 4096;
    int *nbl = malloc(sizeof(int) * maxlen);
    double *tn = malloc(sizeof(double) * nblocks);
    long long nb_total = 0;
    double *tb = malloc(sizeof(double) * nblocks), *ts = malloc(sizeof(double) * nblocks);
    double *tf = malloc(sizeof(double) * nblocks), *tl = malloc(sizeof(double) * nblocks);
    double ti = 0;
    int status = 0;
    for (int a = 3; a < argc && status == 0; ++a) {
        const int ni = atoi(argv[a]);
        if (ni < 1 || ni > npipes || ni > nj) {
            fprintf(stderr, "NI = %d outside [1, min(npipes = %d, NJ)]\n", ni, npipes);
            status = 2;
            break;
        }
        int prev = -1;
        for (int b = -2; b < nblocks; ++b) {   /* two untimed blocks first */
            const int first = (int)(urand() * (nj - ni + 1));
            const double tb0 = now_us();
            if (prev >= 0) {   /* re-send the previous block with its new time */
                for (int k = prev; k < prev + ni; ++k) {
                    double mm = m, d = dtj;
                    int ad = k, idx = k;
                    tj[k] = ti;
                    if (g6_set_j_particle_(&cl, &ad, &idx, &tj[k], &d, &mm, zero, zero, zero,
                                           v[k], x[k])) status = 1;
                }
            }
            ti += 1.0 / 4096;
            g6_set_ti_(&cl, &ti);
            for (int k = 0; k < ni; ++k) {   /* host prediction (first order) */
                const int j = first + k;
                const double dt = ti - tj[j];
                index[k] = j;
                for (int c = 0; c < 3; ++c) {
                    xi[k][c] = x[j][c] + dt * v[j][c];
                    vi[k][c] = v[j][c];
                }
            }
            const double tb1 = now_us();
            double eps2 = 1e-8;
            int njj = nj, nii = ni;
            g6calc_firsthalf_(&cl, &njj, &nii, index, xi, vi, acc, jerk, pot, &eps2, h2);
            const double tb2 = now_us();
            if (g6calc_lasthalf2_(&cl, &njj, &nii, index, xi, vi, &eps2, h2, acc, jerk, pot, nn))
                status = 1;
            const double tb3 = now_us();
            if (h2val > 0) {   /* kira's list retrieval for every i with h2 > 0 */
                if (g6_read_neighbour_list_(&cl) != 0) status = 1;
                for (int k = 0; k < ni; ++k) {
                    int pipe = k, len = 0;
                    if (g6_get_neighbour_list_(&cl, &pipe, &maxlen, &len, nbl) != 0) status = 1;
                    if (b >= 0) nb_total += len;
                }
            }
            const double tb4 = now_us();
            for (int k = 0; k < ni; ++k)
                if (!(pot[k] < 0) || nn[k] < 0 || nn[k] >= nj) status = 1;
            if (dump && b == nblocks - 1 && status == 0 &&
                dump_block(dump, nj, ni, ti, eps2, m, tj, (const double(*)[3])x,
                           (const double(*)[3])v, index, (const double(*)[3])xi,
                           (const double(*)[3])vi, (const double(*)[3])acc,
                           (const double(*)[3])jerk, pot, nn)) {
                fprintf(stderr, "cannot write %s\n", dump);
                status = 1;
            }
            if (b >= 0) {
                tb[b] = tb4 - tb0;
                ts[b] = tb1 - tb0;
                tf[b] = tb2 - tb1;
                tl[b] = tb3 - tb2;
                tn[b] = tb4 - tb3;
            }
            prev = first;
        }
        if (status) {
            fprintf(stderr, "block failed (ni = %d)\n", ni);
            break;
        }
        double sum = 0;
        for (int b = 0; b < nblocks; ++b) sum += tb[b];
        const double mean = sum / nblocks;
        printf("{\"nj\": %d, \"ni\": %d, \"blocks\": %d, \"us_per_block\": %.2f, "
               "\"us_per_block_mean\": %.2f, \"us_set_j_and_predict\": %.2f, "
               "\"us_firsthalf\": %.2f, \"us_lasthalf\": %.2f, \"pairs_per_s\": %.4e, "
               "\"load_us_per_j\": %.3f",
               nj, ni, nblocks, median(tb, nblocks), mean, median(ts, nblocks),
               median(tf, nblocks), median(tl, nblocks), (double)ni * (nj - 1) / (mean * 1e-6),
               load_us / nj);
        if (h2val > 0)
            printf(", \"h2\": %g, \"us_lists\": %.2f, \"mean_list_length\": %.2f",
                   h2val, median(tn, nblocks), (double)nb_total / ((double)nblocks * ni));
        printf("}\n");
        nb_total = 0;
        fflush(stdout);
    }
    g6_close_(&cl);
    return status;
}
