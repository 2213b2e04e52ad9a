/*
 * grape6.h -- the GRAPE-6 C ABI implemented by libg6.so (MI355X / gfx950).
 *
 * These are the entry points that kira's GRAPE-6 driver
 * (src/node/dyn/hdyn/evolve/hdyn_grape6.C) calls.  Names, argument order and
 * meaning are those of the reference declaration include/grape6.h:18-102 so
 * that kira links this library unchanged (Fortran-style trailing underscore,
 * every scalar by pointer, C linkage).  Each declaration below cites the
 * reference line it replaces.  Position/velocity arrays are double[n][3]
 * (the reference passes Starlab `vec` arrays, three contiguous doubles,
 * include/starlab_vector.h:33).
 *
 * Semantics (who calls what is in hdyn_grape6.C):
 *   - j-memory: g6_set_j_particle_ stores one particle at `address`; the
 *     device predicts all j to the time given by g6_set_ti_.
 *   - i-block: g6calc_firsthalf_ starts acc/jerk/pot/nearest-neighbour for
 *     ni i-particles against j-addresses [0, nj); g6calc_lasthalf2_ waits
 *     and returns the results.  j whose index equals index[i] is excluded.
 *   - neighbour lists: g6_read_neighbour_list_ gathers {j : |x_j - x_i|^2 < h2[i]}
 *     for the last block; g6_get_neighbour_list_ copies one i's list.
 *   - return codes: 0 = success.  lasthalf != 0 = device error (kira resets
 *     and retries).  read_neighbour_list > 0 = list storage overflow,
 *     < 0 = error.  get_neighbour_list != 0 = list longer than max_length.
 *   - cluster_id is accepted and ignored (one device per process).
 *   - threads: the library holds one engine per process and is not
 *     re-entrant; call it from one thread at a time, as kira does.
 */
#ifndef LIBG6_MI355X_GRAPE6_H
#define LIBG6_MI355X_GRAPE6_H

#ifdef __cplusplus
extern "C" {
#endif

/* Units of the original fixed-point hardware.  Accepted and ignored: the
 * device computes in floating point.  Reference: grape6.h:18-19. */
void g6_set_tunit_(int *new_tunit);
void g6_set_xunit_(int *new_xunit);

/* Attach / detach the device.  g6_open_ initialises HIP on the current
 * device (or the one named by G6_DEVICE) and allocates nothing until the
 * first j-particle arrives; g6_close_ frees all device memory.  Either may
 * be called at any time; kira reloads every j after re-opening.
 * Reference: grape6.h:21-22. */
int  g6_open_(int *cluster_id);
int  g6_close_(int *cluster_id);

/* Multi-device opens of the sapporo GPU library.  One GPU per process here:
 * the list is accepted, its first entry selects the device.
 * Reference: grape6.h:27-28 (stub variant src/node/dyn/hdyn/stubs/sapporo_stubs.c:18-25). */
int  g6_open_special(int ngpu, int *list);
int  g6_open_special2(int ngpu, int *list, int nmax);

/* Hardware resets.  Both drop pending work and invalidate j-memory
 * (kira reloads it, hdyn_grape6.C:405-438).  Reference: grape6.h:33-34. */
int  g6_reset_(int *cluster_id);
int  g6_reset_fofpga_(int *cluster_id);

/* Largest i-block per call (G6_NPIPES, default 2048).  A value > 48 tells
 * kira it drives a GPU (kira_config.C:108-112), so neighbour lists include
 * the i-particle itself.  Reference: grape6.h:36. */
int  g6_npipes_(void);

/* Time to which the j-particles are predicted in the next calculation.
 * Reference: grape6.h:38. */
void g6_set_ti_(int *cluster_id, double *ti);

/* Store j-particle `address`: tj, dtj (unused), mass, k18 = a''/18,
 * j6 = jerk/6, a2 = acc/2, v, x.  Reference: grape6.h:39-49.
 * A block in flight keeps the j-memory it was started with: a call made
 * between g6calc_firsthalf_ and g6calc_lasthalf2_ (kira makes none,
 * hdyn_grape6.C:675-689) waits until that block's device work is done. */
int  g6_set_j_particle_(int *cluster_id, int *address, int *index,
                        double *tj, double *dtj, double *mass,
                        double k18[3], double j6[3], double a2[3],
                        double v[3], double x[3]);

/* Start the force calculation on ni i-particles (ni <= g6_npipes_()) against
 * j-addresses [0, nj).  aold, j6old, phiold only scaled the old fixed-point
 * pipeline and are ignored.  eps2 is one softening for all i; h2[i] is the
 * neighbour-sphere radius squared (0 = no list).  Reference: grape6.h:55-65. */
void g6calc_firsthalf_(int *cluster_id, int *nj, int *ni, int index[],
                       double xi[][3], double vi[][3], double aold[][3],
                       double j6old[][3], double phiold[], double *eps2,
                       double h2[]);

/* Finish the calculation started by g6calc_firsthalf_: acc, jerk, pot.
 * Reference: grape6.h:67-77. */
int  g6calc_lasthalf_(int *cluster_id, int *nj, int *ni, int index[],
                      double xi[][3], double vi[][3], double *eps2,
                      double h2[], double acc[][3], double jerk[][3],
                      double pot[]);

/* As g6calc_lasthalf_, plus nnbindex[i] = index of the nearest j (smallest
 * |x_j - x_i|^2, lowest address on ties, index[i] excluded; 0 if none).
 * Reference: grape6.h:79-90. */
int  g6calc_lasthalf2_(int *cluster_id, int *nj, int *ni, int index[],
                       double xi[][3], double vi[][3], double *eps2,
                       double h2[], double acc[][3], double jerk[][3],
                       double pot[], int nnbindex[]);

/* Build the neighbour lists of the last calculation.  The _old_ spelling is
 * the same call under G6_OLD_READ_NEIGHBOUR_LIST.  Reference: grape6.h:92-93. */
int  g6_read_neighbour_list_(int *cluster_id);
int  g6_read_neighbour_list_old_(int *cluster_id);

/* Copy the list of i-slot `pipe` (ascending j address) into nbl[0..nblen).
 * Returns 1 and nblen = max_length when the list holds max_length or more
 * entries.  Reference: grape6.h:95-99. */
int  g6_get_neighbour_list_(int *cluster_id, int *pipe, int *max_length,
                            int *nblen, int nbl[]);

/* Host-side j-particle DMA buffering of the PCI GRAPE.  Updates are always
 * buffered and flushed at the next calculation; these are no-ops that
 * return 0.  Reference: grape6.h:101-102. */
int  g6_initialize_jp_buffer_(int *cluster_id, int *size);
int  g6_flush_jp_buffer_(int *cluster_id);

#ifdef __cplusplus
}
#endif
#endif
