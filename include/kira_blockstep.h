/*
 * kira_blockstep.h -- kira's individual block-step Hermite integration as a
 * native caller of a GRAPE-6 ABI library (libkira_blockstep.so).
 *
 * No counterpart in the reference as a library: this is the caller side of
 * the GRAPE-6 boundary, restated for top-level point masses (no tree, no
 * stellar evolution) so that kira's per-block call sequence can be driven
 * at native speed.  The rules it follows are kira's:
 *
 *   block selection  all particles with the smallest t + dt, ascending index
 *                    (fast_get_nodes_to_move, hdyn_schedule.C:641-745)
 *   j update         the previous block re-sent with tj, a/2, j/6, k/18
 *                    (send_j_node_to_grape, hdyn_grape6.C:197-288, :3112-3113)
 *   i prediction     on the host, predict_loworder (include/_dyn_.h:290-295)
 *   force            g6_set_ti_, then g6calc_firsthalf_ / g6calc_lasthalf2_
 *                    in chunks of g6_npipes_() (hdyn_grape6.C:474-910, 3219-3338)
 *   corrector        get_derivs + corrector (hdyn_ev.C:1235-1250, 1717-1718),
 *                    k/18 = bt2 / (9 dt^2) (hdyn_ev.C:1076)
 *   time step        Aarseth criterion, power-of-two rule (hdyn_ev.C:521-887),
 *                    first step set_first_timestep (hdyn_ev.C:377-410)
 *
 * Dot products are x*x + y*y + z*z left to right (Starlab's vec operator*),
 * and the library is compiled without floating-point contraction, so a run
 * is bitwise reproducible and equal to the same rules written elsewhere with
 * the same operation order.
 */
#ifndef KIRA_BLOCKSTEP_H
#define KIRA_BLOCKSTEP_H

#ifdef __cplusplus
extern "C" {
#endif

/* What the library's j-memory holds during a force call: every particle's
 * last sent state (time, mass, x, v, acc, jerk as double[n][3], k18 = the
 * k/18 sent), addresses = particle numbers. */
typedef struct {
    int n;
    const double *t, *m, *x, *v, *acc, *jerk, *k18;
} bs_jmem;

/* Per-block observer, called after each force call with j-memory as above,
 * the block's particles (ascending index), their host-predicted positions and
 * velocities and the library's results.  Return non-zero to stop the run
 * (bs_run then returns -2). */
typedef int (*bs_block_fn)(void *user, const bs_jmem *jmem, double t, int ni,
                           const int *index, const double *xi, const double *vi,
                           const double *acc, const double *jerk, const double *pot,
                           const int *nn);

/* Integrate n point masses from t = 0 to t_end through the GRAPE-6 ABI of
 * the shared library at `grape_lib` (opened with dlopen; it must export
 * g6_open_, g6_close_, g6_npipes_, g6_set_ti_, g6_set_j_particle_,
 * g6calc_firsthalf_ and g6calc_lasthalf2_).  bulk = 1 loads j-particles
 * with g6x_set_j_particles when the library has it (the same records in one
 * call), else one g6_set_j_particle_ per particle as kira does.
 *
 * In: m[n]; x, v as double[n][3] (overwritten with the state at t_end);
 * t_end, eta, dt_max > 0 and eps2 >= 0, all finite (else an error).
 * Out (any may be NULL): acc, jerk as double[n][3]; pot[n]; dt[n] (the last
 * step of each particle); stats[4] = blocks, particle steps, seconds spent
 * inside the ABI calls, seconds in total.
 * Returns 0, -1 on error (bs_last_error), -2 if the observer stopped it. */
int bs_run(const char *grape_lib, int n, const double *m, double *x, double *v,
           double t_end, double eta, double eps2, double dt_max, int bulk,
           double *acc, double *jerk, double *pot, double *dt, double *stats,
           bs_block_fn observer, void *user);

/* ------------------------------------------------------------------------
 * bs_run_kira: the same integration plus the rest of kira's GRAPE-6 call
 * mix (hdyn_grape6.C), each rule restated from the driver:
 *
 *   neighbour radii  set_grape_neighbor_radius (:1851-1985): stars of
 *                    radius R > 0 ask for h2 = max(4 R^2, d_nn^2) every
 *                    coll_freq-th force call; binary CM nodes ask for
 *                    h2 = max(min(gamma^(-2/3) r_bin^2, 100), d_nn^2) every
 *                    pert_freq-th (r_bin = max(a, separation),
 *                    hdyn_inline.C:150-190); nodes with lists last in their
 *                    block (SORT_NODE_LIST, :3118-3147)
 *   lists            g6_read_neighbour_list_ when a chunk has some h2 > 0;
 *                    > 0 (overflow): sort_nodes_and_reduce_rnb halves the
 *                    radii and redoes the chunk (:2192-2255, :3301-3316);
 *                    < 0: reset and redo up to 5 times (:2110-2166);
 *                    g6_get_neighbour_list_ per listed i with max_length
 *                    4096 (:2262-2306): overflow h2 x 0.25, nothing beyond
 *                    the i itself (n_extra = npipes > 48, :2309-2315)
 *                    h2 x 2; either redoes that i alone with h2 x 0.8 per
 *                    further overflow, at most 20 times, and the rest of
 *                    its chunk after it (get_coll_and_perturbers,
 *                    :2803-2998); nn and coll recomputed in FP64 from the
 *                    list (update_nn_coll, hdyn_inline.C:34-82), perturbers
 *                    counted by the hybrid criterion (hdyn_inline.C:289-324)
 *   perturbation     every binary CM node of a block: ni = 2, both slots
 *                    carrying the CM's index, components at the CM's
 *                    predicted position and velocity plus their fixed
 *                    offsets, h2 = 0 (grape6_calculate_perturbation,
 *                    :3415-3587); the offsets do not evolve (the binary's
 *                    internal motion is kira's host-side tree)
 *   energies         at every multiple of dt_log: every node sent with
 *                    t_j = 0 and unpredicted x, pot-only blocks at t_i = 0
 *                    (grape6_calculate_energies, :1403-1546), then a full
 *                    reload before the next block (:3041-3044)
 *   densities        at the same times, if asked: full reload, radii from
 *                    d_nn (set_grape_density_radius, :3607-3685), list
 *                    rounds until each list holds 12 + n_extra stars
 *                    (get_densities / count_neighbors_and_adjust_h2,
 *                    :3704-3898, :4079-4280, :4283-4449)
 *   release          after every release_every blocks g6_close_; the next
 *                    call re-opens and reloads every j (check_release_grape6,
 *                    :958-987; check_reattach_grape6, :3004-3054)
 *   errors           lasthalf2 != 0, or nn < 0 or pot > 0 (error 42),
 *                    resets (g6_reset_, g6_reset_fofpga_, g6_close_,
 *                    g6_open_, full reload) and retries up to 5 times
 *                    (force_by_grape / reset_grape, :405-438, :853-900);
 *                    nn >= nj reads as 0 (:744-754); then the run fails
 *   time offset      t_j and t_i are sent relative to an offset bumped by
 *                    whole units at each full reload (:230-232, :344-345)
 * ------------------------------------------------------------------------ */

/* j-memory as the library holds it: per address, the last record sent
 * (tj and x as sent: t_j = 0 after an energy load). */
typedef struct {
    int n;
    const double *t, *dt, *m, *x, *v, *a2, *j6, *k18;
} bs_jmem_sent;

enum {
    BS_CALL_BLOCK = 0,     /* a chunk of a block's top-level nodes */
    BS_CALL_REDO = 1,      /* one node redone after a list round (get_coll_and_perturbers) */
    BS_CALL_PERTURB = 2,   /* the two components of a binary */
    BS_CALL_ENERGY = 3,    /* pot-only chunk of the energy sweep */
    BS_CALL_DENSITY = 4,   /* a chunk of a density round */
    BS_CALL_LIST = 5       /* one g6_get_neighbour_list_ call */
};
#define BS_NOT_CALLED (-99)

/* One ABI call of bs_run_kira.  Force calls (BLOCK .. DENSITY): ti as sent to
 * g6_set_ti_, nj, ni, index[ni], xi, vi [ni][3], h2[ni] as sent, rc =
 * g6calc_lasthalf2_'s return, the results as returned (undefined when
 * rc != 0), read_status = g6_read_neighbour_list_'s return or BS_NOT_CALLED.
 * LIST: the pipe's index, x and h2 (ni = 1, ti / nj of the force call whose
 * lists it reads), rc = g6_get_neighbour_list_'s return, max_length, nblen and
 * the nblen entries of nbl. */
typedef struct {
    int kind;
    double ti;
    int nj, ni;
    const int *index;
    const double *xi, *vi, *h2;
    int rc;
    const double *acc, *jerk, *pot;
    const int *nn;
    int read_status;
    int pipe, max_length, nblen;
    const int *nbl;
} bs_call;

/* Return non-zero to stop the run (bs_run_kira then returns -2). */
typedef int (*bs_call_fn)(void *user, const bs_jmem_sent *jmem, const bs_call *call);

typedef struct {
    double star_radius;     /* radius R of every single star; 0 = none (no lists for them) */
    int coll_freq;          /* 0 = kira's default 8 (kira_defaults.h:58) */
    int pert_freq;          /* 0 = kira's default 1 (kira_defaults.h:59) */
    double gamma;           /* perturbation threshold; 0 = kira's default 1e-7 (kira_defaults.h:8) */
    int nbin;               /* binary CM nodes: */
    const int *bin_node;    /*   node number of each */
    const double *bin_m1;   /*   mass of component 1 (component 2: the node's mass - m1) */
    const double *bin_dx;   /*   x2 - x1 [nbin][3] */
    const double *bin_dv;   /*   v2 - v1 [nbin][3] */
    double dt_log;          /* energies (and densities) at every multiple of dt_log, t = 0
                               included; 0 = never */
    int densities;          /* 1 = density rounds at the log times */
    int release_every;      /* g6_close_ after every this many blocks; 0 = never */
    double *e_log;          /* out (may be NULL): total energy at each log time, */
    int e_log_cap;          /*   at most this many */
    double *density;        /* out (may be NULL): density[n] of the last density round */
} bs_kira_opts;

/* stats[20]: blocks, particle steps, seconds in ABI calls, seconds in total,
 * force calls, g6_read_neighbour_list_ calls, of them overflows (> 0),
 * g6_get_neighbour_list_ calls, of them overflows, single-node redo calls,
 * perturbation calls, energy sweeps, density rounds, resets, re-opens
 * after a release, full j reloads, list-derived nn updates, perturber lists
 * over 256, error-42 results seen, energy log entries written.
 * The library must also export g6_reset_, g6_reset_fofpga_,
 * g6_read_neighbour_list_ and g6_get_neighbour_list_.  Otherwise as bs_run
 * (no bulk loads: one g6_set_j_particle_ per node, as kira). */
int bs_run_kira(const char *grape_lib, int n, const double *m, double *x, double *v,
                double t_end, double eta, double eps2, double dt_max,
                const bs_kira_opts *opts, double *acc, double *jerk, double *pot,
                double *dt, double *stats, bs_call_fn observer, void *user);

const char *bs_last_error(void);

/* Wall time of the last bs_run by phase, in seconds: seconds[0] the j sends
 * (g6_set_j_particle_ or the bulk call, with the records' packing),
 * [1] g6calc_firsthalf_, [2] g6calc_lasthalf2_ (the wait for the device
 * included), [3] the host prediction of each block's particles, [4] the
 * corrector, new steps and rescheduling, [5] the whole run.  Returns 0, -1
 * for a null array. */
int bs_timers(double *seconds);

#ifdef __cplusplus
}
#endif
#endif
