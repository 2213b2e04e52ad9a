/*
 * g6ens.h -- batched ensemble engine of libg6.so: many independent N-body
 * systems of uneven N integrated together on one MI355X.
 *
 * No counterpart in the reference's C ABI.  The integrator restates kira's
 * Hermite scheme for each system:
 *   block steps      every particle keeps its own power-of-two step; a block
 *                    is every particle with the smallest t + dt
 *                    (hdyn_schedule.C:641-745); all particles are predicted
 *                    to the block time, only the block's are corrected
 *   predictor        include/_dyn_.h:290-295 (predict_loworder)
 *   force            hdyn_ev.C:2019-2036, in FP64 (default) or with libg6's
 *                    FP32 pair law (option "precision")
 *   corrector        hdyn_ev.C:1235-1250 (get_derivs), :1717-1718
 *   step criterion   hdyn_ev.C:627-655 (Aarseth), :855-887 (halve at most
 *                    once, double only when commensurate)
 *   first step       hdyn_ev.C:377-410 (set_first_timestep, eta/16)
 * With option "block_steps" = 0 a system instead takes one shared step, the
 * minimum over its particles of kira's per-particle step, and every particle
 * moves at every step.  Each engine step advances every resident system by
 * one block (or one shared step); systems finish independently at their
 * t_end.  Results of a system are bitwise identical whatever else is
 * resident (its arithmetic depends only on its own particles and N).
 *
 * One engine per process, not re-entrant: call it from one thread at a time.
 *
 * Continuous batching: g6e_submit queues jobs on the host; g6e_run admits
 * queued jobs into free slots while the particle pool (sized from an HBM
 * budget at g6e_open) has room, runs engine steps as HIP-graph replays, and
 * retires finished systems.  Jobs admitted between g6e_run calls overlap on
 * the device with those already running.
 *
 * Conventions: C linkage, double[n][3] AoS vectors, 0 = success, negative =
 * error (text in g6e_last_error()).  One engine per process.
 */
#ifndef LIBG6_MI355X_G6ENS_H
#define LIBG6_MI355X_G6ENS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Reserve device memory for at most `hbm_budget_bytes` of particle state and
 * `max_systems` resident systems.  Opens the device if needed. */
int g6e_open(long long hbm_budget_bytes, int max_systems);
int g6e_close(void);

/* Queue one system (masses m[n], positions x[n][3], velocities v[n][3]) to
 * be integrated from t = 0 to t_end with accuracy parameter eta and
 * softening eps2.  Returns a job id >= 0, or < 0 on error. */
long long g6e_submit(int n, const double *m, const double *x, const double *v,
                     double t_end, double eta, double eps2);

/* Admit queued jobs, run up to `steps` engine steps (in graph replays of
 * the configured length), retire finished systems.  Returns the number of
 * systems resident or queued after the call (0 = idle), < 0 on error. */
int g6e_run(int steps);

/* Job ids finished since the last call (at most max); returns the count. */
int g6e_poll(int max, long long *job_ids);

/* Final state of a finished job: positions, velocities (double[n][3]), and
 * diagnostics: final time, steps taken (blocks in block mode), total energy
 * at t = 0 and at the end
 * (kinetic + potential with the job's softening).  Frees the result. */
int g6e_result(long long job_id, int n, double *x, double *v, double *t_final,
               long long *steps, double *e0, double *e1);

/* Engine statistics since g6e_open: engine steps run, pair interactions
 * evaluated, systems finished, device seconds spent in engine steps. */
int g6e_stats(long long *engine_steps, double *pairs, long long *finished,
              double *device_seconds);

/* Options: "steps_per_graph" (default 8), "jsplit_target" (j per split
 * within a system, default 256; 0 = no split), "max_split" (default 16),
 * "lookahead" (queued jobs examined per admission pass, default 64),
 * "precision" (pair arithmetic of jobs submitted afterwards: 64 = FP64 in the
 * CPU oracle's operation order, the default; 32 = FP32 pairs on
 * double-single positions with a noise-floored step criterion),
 * "block_steps" (time steps of jobs submitted afterwards: 1 = kira's
 * individual block steps, the default; 0 = one shared step per system),
 * "block_fusion" (block-step systems: 1 = prediction inside the force kernel
 * and work lists built on the device, the default; 0 = the four step
 * kernels; fixed while open), "kernel_timing" (1: the step kernels run one
 * at a time between HIP events instead of as graphs, for diagnosis; read
 * the device microseconds per kernel since g6e_open as "kernel_us_predict",
 * "kernel_us_force", "kernel_us_correct", "kernel_us_bselect",
 * "kernel_us_bitems", "kernel_us_bforce", "kernel_us_bcorrect",
 * "kernel_us_finalize", "kernel_us_fused"),
 * "heavy_lanes" (1, the default, or 2: heavy systems alternate between two
 * lanes with their own streams and graphs, so one lane's force kernel can
 * overlap the other's select and correct kernels; faster on large batches
 * when the runtime gives the two streams separate hardware queues, slower
 * when they share one; fixed while open),
 * "max_steps" (per system; beyond it the job fails), "dt_max_bits" (largest
 * block step 2^bits, default -3),
 * "light_max" (systems with n <= light_max run in the light lane, which
 * keeps stepping on its own stream while a heavy-lane graph runs; default
 * 64), "fused_steps" (with light_max <= 64 the light lane runs a fused
 * kernel, one wave per system, up to this many steps per launch; 0 uses
 * step graphs there too; default 256), "wg_steps" (> 0: systems with
 * light_max < n <= wg_max_n run in a workgroup-fused kernel, one workgroup
 * per system (256 threads up to wg_small_max = 512, else 1024), up to this
 * many steps per launch, bitwise equal to the step graphs; shared-step
 * systems only; default 0 = step
 * graphs, which measured faster on mixed cluster batches), "wgb_steps" (> 0:
 * block-step systems of light_max < n <= wg_max_n run in a workgroup kernel,
 * one workgroup per system (256 threads up to wg_small_max, else 512), up to
 * this many block steps per launch, bitwise equal to the block path; default
 * 0: faster on batches of more than a few dozen systems), "wg_small_max",
 * "wg_max_n" (<= 2048); "light_max", "fused_steps", "wg_steps", "wgb_steps", "max_split",
 * "wg_small_max" and "wg_max_n" are fixed while open (set them before
 * g6e_open: they decide the lane each resident system steps in); read-only
 * "medium_steps", and the
 * host time of g6e_run since g6e_open in microseconds by phase:
 * "host_us_retire", "host_us_admit", "host_us_rebuild", "host_us_graphs",
 * "host_us_lanes" (launching and waiting for the lanes), "host_us_phase".
 * Options are library-wide and keep their values across g6e_close.
 * Returns 0, or -1 for an unknown key or an option fixed while open. */
int       g6e_set_option(const char *key, long long value);
long long g6e_get_option(const char *key);

const char *g6e_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
