/*
 * g6_oracle.h -- CPU oracle for the GRAPE-6 force path (TEST INFRASTRUCTURE).
 *
 * It checks results and is never the product: only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  Everything is plain FP64 C,
 * compiled with -ffp-contract=off so every expression rounds exactly as
 * written (no fused multiply-adds), which is what makes bitwise comparison
 * of predicted positions and of nearest-neighbour distances possible.
 *
 * Function -> reference it restates (paths relative to the Starlab tree):
 *   or_ran2_seed / or_randinter   src/std/randinter.C:78-95, 257-324
 *                                 (Numerical Recipes ran2: L'Ecuyer two-LCG
 *                                 combination with a Bays-Durham shuffle)
 *   or_makeplummer                src/node/dyn/init/makeplummer.C:55-214,
 *                                 reshuffle :236-253, seed use :383-395
 *   or_predict                    GRAPE-6 j predictor implied by the j-memory
 *                                 fields of g6_set_j_particle_ (include/grape6.h:39-49,
 *                                 caller src/node/dyn/hdyn/evolve/hdyn_grape6.C:246-249);
 *                                 Taylor series as in include/_dyn_.h:290-295 (4th order)
 *                                 and :312-317 (5th order position term)
 *   or_force                      accumulate_acc_and_jerk, hdyn_ev.C:2019-2036,
 *                                 loop flat_calculate_acc_and_jerk hdyn_ev.C:2047-2227,
 *                                 nn rule update_nn_coll hdyn_inline.C:34-53,
 *                                 exclusion j != index[i] (hdyn_grape6.C:3462-3468)
 *   or_neighbours                 neighbour sphere {j : r^2 < h2_i}
 *                                 (hdyn_grape6.C:2062-2066, 2302-2315)
 *   or_energies                   calculate_energies (src/node/dyn/util, hdyn_tt.C:932-972)
 *   or_density                    compute_density, src/node/dyn/util/compute_density.C:69-226
 *                                 (k-th nearest neighbour density, whole-system mode)
 *   or_g4_force                   GRAPE-4 pair loop of the reference's IEEE simulator,
 *                                 src/harpsims/harp3fastsim.c:240-241, 268-322
 *                                 (acc, jerk, neighbour lists; pot added)
 */
#ifndef G6_ORACLE_H
#define G6_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* --- Starlab random numbers (ran2) --------------------------------------- */
int    or_ran2_seed(int seed);               /* srandinter(seed, 0); returns |seed| */
double or_randinter(double a, double b);     /* a + (b-a)*ran2(randx)               */
int    or_n_rand(void);                      /* number of randinter calls so far    */
int    or_threads(void);                     /* OpenMP threads the oracle will use  */

/* --- initial conditions --------------------------------------------------- */
/* Plummer model exactly as makeplummer -n n -s seed [-u] (mfrac 0.999,
 * reshuffled).  Arrays are n masses, n*3 positions, n*3 velocities (AoS). */
void or_makeplummer(int n, int seed, int scale, double *m, double *x, double *v);

/* --- j predictor ----------------------------------------------------------- */
/* x, v, a2, j6, k18 are nj*3 AoS; tj nj.  order = 4 ignores k18, 5 uses it. */
void or_predict(int nj, const double *x, const double *v, const double *a2,
                const double *j6, const double *k18, const double *tj,
                double ti, int order, double *xp, double *vp);

/* --- direct-sum force ------------------------------------------------------ */
/* For each i: acc, jerk, pot over all j whose jindex != excl[i]; nn = jindex
 * of the j with the strictly smallest d^2 (first in ascending j on ties),
 * dnn2 = that d^2.  nn = -1 and dnn2 = +inf if no j qualifies. */
void or_force(int ni, const double *xi, const double *vi, const int *excl,
              int nj, const double *xj, const double *vj, const double *mj,
              const int *jindex, double eps2,
              double *acc, double *jerk, double *pot, int *nn, double *dnn2);

/* Same, single i, long double accumulation (brute-force reference). */
void or_force_ld(const double *xi, const double *vi, int excl,
                 int nj, const double *xj, const double *vj, const double *mj,
                 const int *jindex, double eps2,
                 double *acc, double *jerk, double *pot);

/* --- neighbour lists ------------------------------------------------------- */
/* For each i with h2[i] > 0: ascending list of jindex with d^2 < h2[i]
 * (d^2 without softening).  include_self = 0 drops jindex == excl[i].
 * counts[i] = true count (may exceed cap); lists[i*cap + k] for k < cap. */
void or_neighbours(int ni, const double *xi, const double *h2, const int *excl,
                   int include_self, int nj, const double *xj, const int *jindex,
                   int cap, int *counts, int *lists);

/* --- GRAPE-4 semantics ------------------------------------------------------ */
/* Per i: r^2 = 2 eps2[i] + d^2 (left to right), pairs with r^2 = 0 skipped,
 * acc, jerk and pot = -sum m/r; counts[i] / lists[i*cap + k] = ascending j
 * addresses with r^2 < 2 h2[i] (counts may exceed cap; counts = NULL skips
 * the lists). */
void or_g4_force(int ni, const double *xi, const double *vi, const double *eps2,
                 const double *h2, int nj, const double *xj, const double *vj,
                 const double *mj, double *acc, double *jerk, double *pot,
                 int cap, int *counts, int *lists);

/* --- densities ------------------------------------------------------------- */
/* For particles i_first .. i_first+ni-1 of the n (x AoS): dk2 = squared
 * distance to the k-th nearest other particle (k < 64), density = masses of
 * the k-1 nearest / (4/3 pi dk2^1.5). */
void or_density(int n, const double *m, const double *x, int i_first, int ni, int k,
                double *density, double *dk2);

/* --- energies ---------------------------------------------------------------*/
/* Pairwise potential (each pair once, softened) and kinetic energy. */
void or_energies(int n, const double *m, const double *x, const double *v,
                 double eps2, double *epot, double *ekin);

/* --- shared-step Hermite integrator (ensemble engine checker) ------------ */
/* One system from t = 0 to t_end with kira's rules: predictor _dyn_.h:290-295,
 * FP64 pair law hdyn_ev.C:2019-2036 (softened), corrector hdyn_ev.C:1235-1250
 * and :1717-1718, Aarseth step :627-655 reduced by min over particles,
 * power-of-two step rule :866-887 (dt <= dt_max), first step :377-410.
 * x, v are updated in place.  Returns 0, or 1 if the step underflowed or
 * max_steps was exceeded.  e0/e1: total energy at start/end. */
int or_hermite_shared(int n, const double *m, double *x, double *v, double t_end,
                      double eta, double eps2, double dt_max, long long max_steps,
                      double *t_final, long long *steps, double *e0, double *e1);

/* --- block-step Hermite integrator (ensemble engine checker, block mode) -- */
/* The same rules with kira's individual power-of-two steps: each block is
 * every particle with the smallest t + dt; all particles are predicted to
 * the block time, the block's forces summed over every other particle in
 * index order, and each block particle corrected and given its own next step
 * (halved once when the criterion falls short, doubled when commensurate,
 * at most min(dt_max, t_end); first steps per particle).  Returns 0, or 1 on
 * step underflow or more than max_blocks blocks.  blocks = block times,
 * psteps = particle steps, t_final = the last block time. */
int or_hermite_block(int n, const double *m, double *x, double *v, double t_end,
                     double eta, double eps2, double dt_max, long long max_blocks,
                     double *t_final, long long *blocks, long long *psteps,
                     double *e0, double *e1);

#ifdef __cplusplus
}
#endif
#endif
