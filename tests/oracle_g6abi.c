/*
 * oracle_g6abi.c -- a GRAPE-6 ABI library whose arithmetic is the FP64 CPU
 * oracle (TEST INFRASTRUCTURE: built by tests only, never shipped).
 *
 * It lets the native block-step driver (libkira_blockstep.so) run on a
 * machine without a GPU, against forces that the Python driver can reproduce
 * exactly: j-memory holds what g6_set_j_particle_ receives, firsthalf
 * predicts every j with or_predict (4th order, the GRAPE form) and sums with
 * or_force; lasthalf2 returns the stored results.  Only the entry points the
 * driver needs are provided.
 */
#include "grape6.h"
#include "g6_oracle.h"

#include <stdlib.h>
#include <string.h>

static int cap, npipes = 64;
static double *X, *V, *A2, *J6, *K18, *T, *M;
static int *IDX;
static double ti_now;
static int last_ni;
static double *r_acc, *r_jerk, *r_pot, *xp, *vp;
static int *r_nn;

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
    const char *e = getenv("ORACLE_G6ABI_NPIPES");
    if (e) npipes = atoi(e);
    return 0;
}

int g6_close_(int *c)
{
    (void)c;
    return 0;
}

int g6_npipes_(void) { return npipes; }

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

void g6calc_firsthalf_(int *c, int *nj, int *ni, int index[], double xi[][3],
                       double vi[][3], double aold[][3], double j6old[][3],
                       double phiold[], double *eps2, double h2[])
{
    (void)c; (void)aold; (void)j6old; (void)phiold; (void)h2;
    const int n = *nj, k = *ni;
    grow(n > k ? n : k);
    or_predict(n, X, V, A2, J6, K18, T, ti_now, 4, xp, vp);
    r_acc = realloc(r_acc, sizeof(double) * 3 * k);
    r_jerk = realloc(r_jerk, sizeof(double) * 3 * k);
    r_pot = realloc(r_pot, sizeof(double) * k);
    r_nn = realloc(r_nn, sizeof(int) * k);
    double *d2 = malloc(sizeof(double) * k);
    or_force(k, &xi[0][0], &vi[0][0], index, n, xp, vp, M, IDX, *eps2, r_acc, r_jerk, r_pot,
             r_nn, d2);
    free(d2);
    last_ni = k;
}

int g6calc_lasthalf2_(int *c, int *nj, int *ni, int index[], double xi[][3], double vi[][3],
                      double *eps2, double h2[], double acc[][3], double jerk[][3],
                      double pot[], int nnbindex[])
{
    (void)c; (void)nj; (void)index; (void)xi; (void)vi; (void)eps2; (void)h2;
    if (*ni != last_ni) return 1;
    memcpy(&acc[0][0], r_acc, sizeof(double) * 3 * last_ni);
    memcpy(&jerk[0][0], r_jerk, sizeof(double) * 3 * last_ni);
    memcpy(pot, r_pot, sizeof(double) * last_ni);
    memcpy(nnbindex, r_nn, sizeof(int) * last_ni);
    return 0;
}
