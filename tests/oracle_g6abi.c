/*
 * oracle_g6abi.c -- a GRAPE-6 ABI library whose arithmetic is the FP64 CPU
 * oracle (TEST INFRASTRUCTURE: built by tests only, never shipped).
 *
 * It lets the native block-step driver (libkira_blockstep.so) run on a
 * machine without a GPU, against forces that the Python driver can reproduce
 * exactly: j-memory holds what g6_set_j_particle_ receives, firsthalf
 * predicts every j with or_predict (4th order, the GRAPE form) and sums with
 * or_force; lasthalf2 returns the stored results.  Neighbour lists are the
 * brute-force FP64 sets of or_neighbours with libg6's conventions (the i
 * itself included when npipes > 48, at most NB_CAP entries stored per i
 * (8192 as libg6's nb_cap, or ORACLE_G6ABI_NB_CAP);
 * read returns 1 when a list exceeds that, get returns 1 with nblen =
 * max_length when the list has max_length or more entries).  Resets and
 * re-opens keep nothing on the "device" side that the driver relies on.
 * Failure injection for the driver's recovery path, read from the
 * environment whenever a variable's text has changed since the last open: ORACLE_G6ABI_FAIL_LASTHALF=k makes the next k lasthalf2 calls
 * return 1, ORACLE_G6ABI_FAIL_READNL=k the next k read_neighbour_list calls
 * return -1, ORACLE_G6ABI_FAIL_NN=k the next k lasthalf2 calls report
 * nn = -1 for slot 0 (kira's error 42).  ORACLE_G6ABI_JSPLIT=T (T > 1)
 * sums blocks of fewer than 64 i over T j-ranges in parallel and adds the
 * ranges in order: for long runs at large N, where one i per thread leaves
 * most threads idle; the sums then differ from or_force's in the last bits,
 * the nearest neighbour (smallest d2, lowest j on ties) does not.  Only the
 * entry points the driver needs are provided.
 */
#include "grape6.h"
#include "g6_oracle.h"

#include <stdlib.h>
#include <string.h>

static int cap, npipes = 64;
static double *X, *V, *A2, *J6, *K18, *T, *M;
static int *IDX;
static double ti_now;
static int last_ni, last_nj;
static double *r_acc, *r_jerk, *r_pot, *xp, *vp;
static int *r_nn;
static double *xi_now, *h2_now;
static int *excl_now;
static int NB_CAP = 8192;   /* ORACLE_G6ABI_NB_CAP */
static int *nb_counts, *nb_lists;
static int nb_valid, fail_lasthalf, fail_readnl, fail_nn, env_read;
static int jsplit = 1;   /* ORACLE_G6ABI_JSPLIT */

/* A setting is taken when its variable's text differs from the text seen
 * last (the library stays loaded across driver runs, and the re-opens of a
 * reset must not re-arm a failure count). */
static int env_changed(const char *name, char *last, int *value, int dflt)
{
    const char *e = getenv(name);
    const char *txt = e ? e : "";
    if (strncmp(txt, last, 63) == 0) return 0;
    strncpy(last, txt, 63);
    last[63] = 0;
    *value = e ? atoi(e) : dflt;
    return 1;
}

static void read_env(void)
{
    static char s_np[64], s_fl[64], s_fr[64], s_fn[64], s_nb[64], s_js[64];
    if (!env_read) {   /* first call: "" differs from an unset-variable marker */
        strcpy(s_np, "\x01"); strcpy(s_fl, "\x01"); strcpy(s_fr, "\x01");
        strcpy(s_fn, "\x01"); strcpy(s_nb, "\x01"); strcpy(s_js, "\x01");
        env_read = 1;
    }
    if (env_changed("ORACLE_G6ABI_JSPLIT", s_js, &jsplit, 1) && jsplit < 1) jsplit = 1;
    env_changed("ORACLE_G6ABI_NPIPES", s_np, &npipes, 64);
    env_changed("ORACLE_G6ABI_FAIL_LASTHALF", s_fl, &fail_lasthalf, 0);
    env_changed("ORACLE_G6ABI_FAIL_READNL", s_fr, &fail_readnl, 0);
    env_changed("ORACLE_G6ABI_FAIL_NN", s_fn, &fail_nn, 0);
    if (env_changed("ORACLE_G6ABI_NB_CAP", s_nb, &NB_CAP, 8192) && NB_CAP < 1) NB_CAP = 8192;
}

static void grow(int need)
{
    if (need <= cap) return;
    int c = cap ? cap : 256;
    while (c < need) c *= 2;
    X = realloc(X, sizeof(double) * 3 * c);
    V = realloc(V, sizeof(double) * 3 * c);
    A2 = realloc(A2, sizeof(double) * 3 * c);
    J6 = realloc(J6, sizeof(double) * 3 * c);
    K18 = realloc(K18, sizeof(double) * 3 * c);
    xp = realloc(xp, sizeof(double) * 3 * c);
    vp = realloc(vp, sizeof(double) * 3 * c);
    T = realloc(T, sizeof(double) * c);
    M = realloc(M, sizeof(double) * c);
    IDX = realloc(IDX, sizeof(int) * c);
    for (int k = cap; k < c; ++k) IDX[k] = -1;
    cap = c;
}

int g6_open_(int *c)
{
    (void)c;
    read_env();
    return 0;
}

int g6_close_(int *c)
{
    (void)c;
    nb_valid = 0;
    return 0;
}

int g6_reset_(int *c)
{
    (void)c;
    nb_valid = 0;
    return 0;
}

int g6_reset_fofpga_(int *c) { return g6_reset_(c); }

int g6_npipes_(void)
{
    read_env();
    return npipes;
}

void g6_set_ti_(int *c, double *ti)
{
    (void)c;
    ti_now = *ti;
}

int g6_set_j_particle_(int *c, int *address, int *index, double *tj, double *dtj,
                       double *mass, double k18[3], double j6[3], double a2[3],
                       double v[3], double x[3])
{
    (void)c;
    (void)dtj;
    const int a = *address;
    if (a < 0) return -1;
    grow(a + 1);
    memcpy(X + 3 * a, x, sizeof(double) * 3);
    memcpy(V + 3 * a, v, sizeof(double) * 3);
    memcpy(A2 + 3 * a, a2, sizeof(double) * 3);
    memcpy(J6 + 3 * a, j6, sizeof(double) * 3);
    memcpy(K18 + 3 * a, k18, sizeof(double) * 3);
    T[a] = *tj;
    M[a] = *mass;
    IDX[a] = *index;
    return 0;
}

/* or_force over jsplit j-ranges at once, the ranges added in order (the
 * ORACLE_G6ABI_JSPLIT mode); results into r_acc, r_jerk, r_pot, r_nn. */
static void force_jsplit(int k, const double *xi, const double *vi, const int *excl, int n,
                         double eps2)
{
    const int T = jsplit, chunk = (n + T - 1) / T;
    double *pa = malloc(sizeof(double) * 3 * k * T), *pj = malloc(sizeof(double) * 3 * k * T);
    double *pp = malloc(sizeof(double) * k * T), *pd = malloc(sizeof(double) * k * T);
    int *pn = malloc(sizeof(int) * k * T);
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (int c = 0; c < T; ++c) {
        const int j0 = c * chunk, nc = j0 < n ? (n - j0 < chunk ? n - j0 : chunk) : 0;
        or_force(k, xi, vi, excl, nc, xp + 3 * (size_t)j0, vp + 3 * (size_t)j0, M + j0,
                 IDX + j0, eps2, pa + 3 * k * c, pj + 3 * k * c, pp + k * c, pn + k * c,
                 pd + k * c);
    }
    for (int q = 0; q < k; ++q) {
        double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, p = 0, best = pd[q];
        int nn = pn[q];
        for (int c = 0; c < T; ++c) {
            for (int d = 0; d < 3; ++d) {
                a[d] += pa[3 * (k * c + q) + d];
                jk[d] += pj[3 * (k * c + q) + d];
            }
            p += pp[k * c + q];
            if (c > 0 && pd[k * c + q] < best) {   /* strict <: the lower range wins ties */
                best = pd[k * c + q];
                nn = pn[k * c + q];
            }
        }
        memcpy(r_acc + 3 * q, a, sizeof a);
        memcpy(r_jerk + 3 * q, jk, sizeof jk);
        r_pot[q] = p;
        r_nn[q] = nn;
    }
    free(pa); free(pj); free(pp); free(pd); free(pn);
}

void g6calc_firsthalf_(int *c, int *nj, int *ni, int index[], double xi[][3],
                       double vi[][3], double aold[][3], double j6old[][3],
                       double phiold[], double *eps2, double h2[])
{
    (void)c; (void)aold; (void)j6old; (void)phiold;
    const int n = *nj, k = *ni;
    grow(n > k ? n : k);
    or_predict(n, X, V, A2, J6, K18, T, ti_now, 4, xp, vp);
    r_acc = realloc(r_acc, sizeof(double) * 3 * k);
    r_jerk = realloc(r_jerk, sizeof(double) * 3 * k);
    r_pot = realloc(r_pot, sizeof(double) * k);
    r_nn = realloc(r_nn, sizeof(int) * k);
    double *d2 = malloc(sizeof(double) * k);
    if (jsplit > 1 && k < 64 && n >= 64 * jsplit)
        force_jsplit(k, &xi[0][0], &vi[0][0], index, n, *eps2);
    else
        or_force(k, &xi[0][0], &vi[0][0], index, n, xp, vp, M, IDX, *eps2, r_acc, r_jerk,
                 r_pot, r_nn, d2);
    free(d2);
    last_ni = k;
    last_nj = n;
    xi_now = realloc(xi_now, sizeof(double) * 3 * k);
    h2_now = realloc(h2_now, sizeof(double) * k);
    excl_now = realloc(excl_now, sizeof(int) * k);
    memcpy(xi_now, &xi[0][0], sizeof(double) * 3 * k);
    memcpy(excl_now, index, sizeof(int) * k);
    for (int q = 0; q < k; ++q) h2_now[q] = h2 ? h2[q] : 0.0;
    nb_valid = 0;
}

int g6calc_lasthalf2_(int *c, int *nj, int *ni, int index[], double xi[][3], double vi[][3],
                      double *eps2, double h2[], double acc[][3], double jerk[][3],
                      double pot[], int nnbindex[])
{
    (void)c; (void)nj; (void)index; (void)xi; (void)vi; (void)eps2; (void)h2;
    if (*ni != last_ni) return 1;
    if (fail_lasthalf > 0) {
        --fail_lasthalf;
        return 1;
    }
    memcpy(&acc[0][0], r_acc, sizeof(double) * 3 * last_ni);
    memcpy(&jerk[0][0], r_jerk, sizeof(double) * 3 * last_ni);
    memcpy(pot, r_pot, sizeof(double) * last_ni);
    memcpy(nnbindex, r_nn, sizeof(int) * last_ni);
    if (fail_nn > 0 && last_ni > 0) {
        --fail_nn;
        nnbindex[0] = -1;
    }
    return 0;
}

int g6_read_neighbour_list_(int *c)
{
    (void)c;
    if (fail_readnl > 0) {
        --fail_readnl;
        return -1;
    }
    const int k = last_ni;
    int *ex = malloc(sizeof(int) * (k > 0 ? k : 1));
    for (int q = 0; q < k; ++q) {   /* the excluded address: the j whose index field is excl */
        ex[q] = excl_now[q];
    }
    nb_counts = realloc(nb_counts, sizeof(int) * (k > 0 ? k : 1));
    nb_lists = realloc(nb_lists, sizeof(int) * (size_t)NB_CAP * (k > 0 ? k : 1));
    or_neighbours(k, xi_now, h2_now, ex, npipes > 48, last_nj, xp, IDX, NB_CAP, nb_counts,
                  nb_lists);
    free(ex);
    nb_valid = 1;
    for (int q = 0; q < k; ++q)
        if (nb_counts[q] > NB_CAP) return 1;
    return 0;
}

int g6_get_neighbour_list_(int *c, int *pipe, int *max_length, int *nblen, int nbl[])
{
    (void)c;
    if (!nb_valid || *pipe < 0 || *pipe >= last_ni) {
        *nblen = 0;
        return -1;
    }
    const int cnt = nb_counts[*pipe];
    const int stored = cnt < NB_CAP ? cnt : NB_CAP;
    const int n = stored < *max_length ? stored : *max_length;
    if (n > 0) memcpy(nbl, nb_lists + (size_t)NB_CAP * *pipe, sizeof(int) * n);
    if (cnt >= *max_length) {
        *nblen = *max_length;
        return 1;
    }
    *nblen = cnt;
    return cnt > NB_CAP ? 1 : 0;
}
