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

const char *bs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
