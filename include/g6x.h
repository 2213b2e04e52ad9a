/*
 * g6x.h -- libg6.so extensions beyond the GRAPE-6 ABI.
 *
 * None of these has a counterpart in the reference (kira calls only the
 * grape6.h entry points).  They exist for callers that hold many particles
 * at once -- benchmarks, test harnesses, ensemble drivers -- and for
 * measurement.  All follow the ABI's conventions: C linkage, plain pointers,
 * double[n][3] AoS vectors, 0 = success, negative = error (see
 * g6x_last_error()).
 */
#ifndef LIBG6_MI355X_G6X_H
#define LIBG6_MI355X_G6X_H

#ifdef __cplusplus
extern "C" {
#endif

/* Bulk form of g6_set_j_particle_: n particles, address[k] / index[k],
 * tj[k], dtj[k], mass[k], and k18, j6, a2, v, x as double[n][3].  dtj may
 * be NULL (1.0 each); any other NULL array is an error. */
int g6x_set_j_particles(int cluster_id, int n, const int *address,
                        const int *index, const double *tj, const double *dtj,
                        const double *mass, const double *k18, const double *j6,
                        const double *a2, const double *v, const double *x);

/* Device-resident sweep: the j-particles at addresses [i_first, i_first+ni)
 * are themselves the i-particles, at their device-predicted positions and
 * velocities for time ti (each excludes itself).  Runs predictor + force +
 * reduction + nearest-neighbour refinement on the device and leaves the
 * results there; nothing crosses PCIe.  Asynchronous: g6x_sync() waits. */
int g6x_self_sweep(int cluster_id, int nj, int i_first, int ni, double ti,
                   double eps2);

/* Potential only (the fast path for kira's energy computation,
 * grape6_calculate_energies, hdyn_grape6.C:1403-1546, which reads only pot
 * from its pot_only force calls, :477-551, :766-767): pot[k] = -sum_j m_j/r
 * over j in [0, nj) for the j-particle at address i_first + k, all at their
 * device-predicted positions for time ti (ti = 0 with tj = 0 is kira's
 * unpredicted energy snapshot), itself excluded, softened with eps2.
 * Same double-single / FP32 pair arithmetic and FP64 sums as the force
 * path, about 2.5x fewer instructions per pair.  Synchronous. */
int g6x_potential(int cluster_id, int nj, int i_first, int ni, double ti, double eps2,
                  double *pot);

/* Local densities (the fast path for grape6_calculate_densities,
 * hdyn_grape6.C:4283-4449, which iterates neighbour radii until each list
 * holds 12 stars and then calls compute_density,
 * src/node/dyn/util/compute_density.C:69-226): for the j-particle at
 * address i_first + q, dk2[q] = squared distance to its k-th nearest other
 * j (FP64, exact) and density[q] = (sum of the masses of its k-1 nearest) /
 * (4/3 pi dk2^1.5), 0 when dk2 = 0.  2 <= k <= 32 (kira uses 12), nj > k.
 * Either output may be NULL.  Synchronous. */
int g6x_densities(int cluster_id, int nj, int i_first, int ni, double ti, int k,
                  double *density, double *dk2);

/* g6x_self_sweep for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The same computation, geometry and
 * results as g6x_self_sweep, but its scratch -- the i-block, the j-split
 * partial sums and the refinement flags -- lives in a fixed 48 MiB device
 * array per GPU (held until the process ends) instead of library
 * allocations, and the results go straight to the caller's DEVICE arrays
 * acc[ni][3], jerk[ni][3], pot[ni], nnbindex[ni].  It calls no hipMalloc,
 * hipFree or other allocating routine, and event timing is off for the call.
 * Preconditions, each refused with an error (g6x_last_error) instead of
 * growing a buffer: j-memory holds at least nj addresses and is resident on
 * the device (any force call, g6x_self_sweep or g6x_autotune since the last
 * new address places it), and the geometry for (ni, nj) fits the array (the
 * default geometry, waves_target 16384, fits every ni up to 65536 at any
 * nj: at most 36.5 MiB).  j updates made
 * since then go up by copy or by the existing staging buffer.
 * Asynchronous on the library stream: g6x_sync() waits.  Every call on a GPU
 * shares that GPU's array and keeps the previous call's contents (nothing in
 * it needs clearing); calls are ordered on the library stream, so only
 * callers that enter the library from several threads at once race.
 * g6x_read_results does not return these results. */
int g6x_self_sweep_noalloc(int cluster_id, int nj, int i_first, int ni, double ti,
                           double eps2, double *acc, double *jerk, double *pot,
                           int *nnbindex);

/* Copy the results of the last g6x_self_sweep (ni entries) to the host.
 * Returns -1 if the last calculation was anything else (an ABI block,
 * g6x_potential, g6x_densities, g6x_autotune), if that sweep failed, or if a
 * close came after it. */
int g6x_read_results(int cluster_id, int ni, double *acc, double *jerk,
                     double *pot, int *nnbindex);

/* Device-predicted j state at the last prediction (double[nj][3] each). */
int g6x_read_predicted(int cluster_id, int nj, double *xp, double *vp);

/* Wait for all queued device work. */
int g6x_sync(int cluster_id);

/* Options (also read from the environment at g6_open_):
 *   "predict_order"  4 (default: a2, j6) or 5 (adds k18)     G6_PREDICT_ORDER
 *   "npipes"         value returned by g6_npipes_            G6_NPIPES
 *   "nb_cap"         neighbour entries stored per i          G6_NB_CAP
 *   "waves_target"   waves the force grid aims for           G6_WAVES
 *   "nsplit"         force j-split (0 = automatic); also the j-split of the
 *                    k-nearest-neighbour sweep of g6x_densities   G6_NSPLIT
 *   "timing"         1 = time each kernel with HIP events    G6_TIMING
 *   "force_variant"  -1 auto, 0 scalar-load, 1 LDS-staged     G6_FORCE_VARIANT
 *   "lds_min_ni"     auto picks the LDS kernel from this ni  G6_LDS_MIN_NI
 *   "xcd_remap"      1 = keep each j-split on one XCD         G6_XCD_REMAP
 *   "tile"           j per FP32 tile: 0 auto, 16, 32, 64, 128, 256   G6_TILE
 *                    (16 and 32 suit blocks of one or a few i-groups)
 *   "last_tile"      (read only) tile of the last force launch
 *   "reduce_lanes"   lanes per i in the split reduction: 0 auto, 4..64  G6_REDUCE_LANES
 *   "wpg"            waves per force workgroup, each on its own j-split, folded
 *                    in LDS before the reduction: 0 auto, 1, 2, 4, 8, 16  G6_WPG
 *   "last_wpg", "last_lanes"  (read only) of the last force launch
 *   "fused_small"    1 (default) = ABI blocks of at most 256 i with one
 *                    softening predict j inside the force kernel (no
 *                    predictor launch, no JPred round trip; bitwise the
 *                    same results); 0 = predictor first   G6_FUSED_SMALL
 *   "small_fp64"     1 (default) = ABI blocks of at most 12 i with one
 *                    softening run as one launch: FP64 pair arithmetic with
 *                    the j-particles across lanes, the reduction in the same
 *                    kernel, up to 16 pending j updates carried in the launch
 *                    arguments; results within ~1e-15 of the FP64 direct sum,
 *                    nearest neighbours exact.  0 = the FP32 path
 *                    G6_SMALL_FP64
 *   "done_word"      1 (default) = the workgroup that finishes an ABI
 *                    block's last launch (the small-block kernel, or the
 *                    split reduction of a block of at most 512 i) stores a
 *                    sequence number in a pinned word after the results,
 *                    and lasthalf polls it (at most ~2 ms, then a stream
 *                    wait) instead of waiting for an event's completion
 *                    signal; 0 = the event (also while "timing" is on)
 *                    G6_DONE_WORD
 *   "fail_lasthalf", "fail_read_nl", "fail_nn"   host-side failure injection
 *                    for a caller's recovery path (kira resets, reloads and
 *                    retries, hdyn_grape6.C:405-438, 853-900, 2110-2166):
 *                    the next k g6calc_lasthalf(2)_ calls return 1 (results
 *                    withheld), the next k g6_read_neighbour_list_ calls
 *                    return -1, the next k g6calc_lasthalf2_ calls return 0
 *                    with nnbindex[0] = -1 (kira's error 42).  The device
 *                    work runs normally; each counts down to 0.
 *                    G6_FAIL_LASTHALF, G6_FAIL_READNL, G6_FAIL_NN
 * Returns 0, or -1 for an unknown key. */
int       g6x_set_option(const char *key, long long value);
long long g6x_get_option(const char *key);

/* Per-kernel event timing, accumulated since the last reset (needs
 * "timing" = 1).  kernel: 0 predictor, 1 force (with the prediction for
 * fused small blocks, and the whole calculation of a "small_fp64" block),
 * 2 reduce (includes the
 * FP64 nearest-neighbour rescan of ambiguous i), 3 unused (0 launches),
 * 4 neighbour list, 5 j scatter, 6 potential (g6x_potential), 7 k-nearest
 * neighbours (g6x_densities).  Returns launches; *ms = total ms.  Kernel 4
 * covers the four launches of one neighbour-list build. */
int  g6x_kernel_time(int kernel, double *ms);
void g6x_reset_timing(void);

/* Geometry of the last force launch: workgroups, j-split, j per split.
 * Returns the kernel variant (0: scalar-load kernel, 1: LDS-staged,
 * 2: the one-launch FP64 kernel of small blocks, whose j-split is its
 * workgroup count and whose j per split is 256 x j per lane). */
int  g6x_last_force_geometry(int *workgroups, int *nsplit, int *jchunk);

/* Number of nearest-neighbour refinements (FP64 rescans) in the last call. */
int  g6x_last_refinements(void);

/* Autotuner.  For each i-block size bucket (bucket b: 2^(b-1) < ni <= 2^b)
 * up to ni_max, times force + reduction for every candidate shape (kernel
 * variant, j-tile length 64/128/256 and, for ni <= 64, 16/32, waves per
 * launch 2048..32768, waves per workgroup, then the reduction's lanes per i)
 * against the first nj j-particles as loaded, and keeps the fastest.  Later force calls of that size use the tuned shape
 * unless an explicit option overrides it.  Takes ~1 s at nj = 65536.
 * Returns the number of buckets tuned, or -1. */
int  g6x_autotune(int cluster_id, int nj, int ni_max, double eps2);
/* Tuned shape of one bucket; returns 1 if tuned, 0 if not, -1 bad bucket. */
int  g6x_get_tuning(int bucket, int *variant, int *tile, int *waves, int *lanes);
int  g6x_clear_tuning(void);
/* Tuning table as text ("bucket variant tile waves lanes wpg" per line; a
 * line without wpg uses the built-in rule for it).  The
 * library also loads the file named by G6_TUNE_FILE at g6_open_.
 * save returns 0 / -1; load returns the entries read or -1. */
int  g6x_save_tuning(const char *path);
int  g6x_load_tuning(const char *path);

/* Text of the last error (empty if none). */
const char *g6x_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
