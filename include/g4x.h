/*
 * g4x.h -- libg4x.so extensions beyond the GRAPE-4 ABI.
 *
 * None of these has a counterpart in the reference.  The GRAPE-4 entry
 * points return no status, so a harness that runs with G4_EXIT_ON_ERROR=0
 * reads failures here; the options let tests and tools set what the
 * environment would otherwise set.
 */
#ifndef LIBG6_MI355X_G4X_H
#define LIBG6_MI355X_G4X_H

#ifdef __cplusplus
extern "C" {
#endif

/* Number of failed GRAPE-4 calls since the library was loaded. */
long long g4x_errors(void);

/* Text of the last error ("" if none). */
const char *g4x_last_error(void);

/* Options: "npipes", "jpmax", "nb_cap" (list entries stored per pipe before
 * the store grows), "exit_on_error", "dedupe" (compute identical adjacent
 * pipes once), "timing", "fused_small" (1, the default: blocks of at most
 * 256 distinct i with one softening predict j inside the force kernel, with
 * the same results); read-only "last_slots" (distinct i-particles of
 * the last block) and "last_groups" (softening groups).  set returns 0 or
 * -1 for an unknown key; get returns -1 for an unknown key. */
int g4x_set_option(const char *key, long long value);
long long g4x_get_option(const char *key);

#ifdef __cplusplus
}
#endif
#endif
