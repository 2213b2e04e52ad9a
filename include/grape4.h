/*
 * grape4.h -- the GRAPE-4 (HARP-3) C ABI implemented by libg4x.so
 * (MI355X / gfx950), on the same kernels as libg6.so.
 *
 * These are the entry points that kira's GRAPE-4 driver
 * (src/node/dyn/hdyn/evolve/hdyn_grape4.C) calls.  Names, argument order and
 * meaning are those of the reference declaration include/grape4.h:22-45, so
 * kira links this library unchanged; the library is named libg4x because
 * Starlab's configure probes exactly that name for h3open_
 * (configure.ac:214-221).  Each declaration cites the reference line it
 * replaces.  Position-like arrays are double[n][3] (the reference passes
 * Starlab `vec` arrays, three contiguous doubles, include/starlab_vector.h:33);
 * `real` is double (include/stdinc.h:130).
 *
 * Semantics (those of the reference's IEEE GRAPE-4 simulator,
 * src/harpsims/harp3fastsim.c):
 *   - j-memory: a DMA call copies nj particles, named by 1-based hostindex
 *     into the caller's arrays, into a buffer; starting the buffer writes
 *     them to j-memory address hostindex-1.  aj = acc/2, jj = jerk/6.
 *   - i-block: h3calc_firsthalf_ starts acc/jerk/pot for ni pipes against
 *     j-addresses [0, nj) with every j predicted to the time of h3setti_;
 *     h3calc_lasthalf_ waits and returns them.  eps2 and h2 are per pipe and
 *     doubled internally: r^2 = 2 eps2 + |x_j - x_i|^2; pairs with r^2 = 0
 *     (the i itself when eps2 = 0) add nothing.
 *   - neighbour lists: h3nbread_ builds {j : r^2 < 2 h2} for every pipe of
 *     the last block, the i itself included; h3nblist_ copies one pipe's
 *     list in ascending address and returns its length.  Lists are always
 *     complete: the caller's buffer must hold up to nj entries.
 *   - errors: these calls return no status.  A failure prints a message and
 *     exits the process unless G4_EXIT_ON_ERROR=0 (then see g4x.h).
 */
#ifndef LIBG6_MI355X_GRAPE4_H
#define LIBG6_MI355X_GRAPE4_H

#ifdef __cplusplus
extern "C" {
#endif

/* Attach / detach the device (grape4.h:22-23).  h3close_ frees all device
 * memory; j-memory contents survive in host memory and return on the next
 * h3open_. */
void h3open_(void);
void h3close_(void);

/* Number of i pipes per h3calc call (grape4.h:24); G4_NPIPES, default 2048. */
int h3npipe_(void);

/* Board count (grape4.h:25-26): one device; setting is accepted and ignored. */
void h3setnboards_(int *nboards);
int h3getnboards_(void);

/* Wait for the board (grape4.h:27).  Calls are ordered on one device stream,
 * so there is nothing to wait for here; returns 0. */
unsigned int h3wait_(void);

/* Time to which all j are predicted by the next h3calc (grape4.h:28). */
void h3setti_(double *ti);

/* Board mode and LEDs (grape4.h:29-30): accepted and ignored. */
void h3setmode_(int *mode, int *iboard);
void h3setled_(int *mode, int *iboard);

/* Largest nj the caller should put in one DMA call (grape4.h:31);
 * G4_JPMAX, default 4096.  Larger calls are accepted. */
int h3jpmax_(void);

/* Copy nj particles into DMA buffer 0 and write it to j-memory
 * (grape4.h:32).  mode is accepted and ignored. */
void h3jpdma_indirect_(int *nj, int *hostindex, double xj[][3], double vj[][3],
                       double aj[][3], double jj[][3], double *mj, double *tj,
                       int *mode);

/* Copy nj particles into DMA buffer *buff_id (the next free buffer when
 * *buff_id < 0; at most 1024 buffers), to be written by h3mjpdma_start_ or
 * h3mjpdma_flush_ (grape4.h:33). */
void h3mjpdma_indirect_(int *nj, int *hostindex, double xj[][3], double vj[][3],
                        double aj[][3], double jj[][3], double *mj, double *tj,
                        int *mode, int *buff_id);

/* Write buffer *buff_id to j-memory (grape4.h:34). */
void h3mjpdma_start_(int *buff_id);

/* Write every buffer taken with buff_id < 0 (grape4.h:35). */
void h3mjpdma_flush_(void);

/* firsthalf + lasthalf in one call (grape4.h:36-37). */
void h3calc_(int *nj, int *ni, double xi[][3], double vi[][3], double *eps2,
             double *h2, double acc[][3], double jerk[][3], double *pot);

/* Start the force on ni <= h3npipe_() pipes (grape4.h:38).  Asynchronous. */
void h3calc_firsthalf_(int *nj, int *ni, double xi[][3], double vi[][3],
                       double *eps2, double *h2);

/* Wait and return acc, jerk, pot of the pipes started last (grape4.h:39). */
void h3calc_lasthalf_(int *ni, double acc[][3], double jerk[][3], double *pot);

/* Build the neighbour lists of the last i-block (grape4.h:40). */
void h3nbread_(int *nboards);

/* Copy the list of pipe *chip into nblist; returns its length
 * (grape4.h:41).  board is ignored. */
int h3nblist_(int *board, int *chip, int *nblist);

/* Length of the list of pipe chip, 0 before h3nbread_ (grape4.h:42). */
int count_nblist_low(int board, int chip);

/* Debug level, DMA-buffer check and time-check mode (grape4.h:43-45):
 * accepted and ignored. */
void h3setdebuglevel_(int *debug_level);
void h3checkdmabuffers_(void);
void set_time_check_mode(int mode);

#ifdef __cplusplus
}
#endif
#endif
