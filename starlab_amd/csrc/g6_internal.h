// g6_internal.h -- device data layout and kernel launchers of libg6 (gfx950).
//
// HBM layout (one process = one device = one j-memory):
//   JRaw   jraw[cap]        160 B/j  as stored by g6_set_j_particle_ (FP64)
//   JPred  jpred[cap]        64 B/j  predicted: double-single position,
//                                    FP32 velocity and mass (force kernel)
//   double xp64[3][cap]      24 B/j  predicted FP64 position (nn refinement,
//                                    neighbour lists; bit-exact distances)
//   double vp64[3][cap]      24 B/j  predicted FP64 velocity (read-back, FP64 nearest tile)
//   double m64[cap]           8 B/j  FP64 mass, right after vp64 (FP64 nearest tile);
//                                    written with every prediction
//   int    jindex[cap]        4 B/j  the `index` field of each address
// i-block (npipes slots, SoA so each lane reads its own slot coalesced):
//   10 arrays of float/int, FP64 positions 3 arrays, h2 and the list offset
//   e2, FP64 velocities 3 arrays
// Partial sums of the j-split force: 7 FP64 arrays + 3 nn words + the 10
// FP32 sums of the split's nearest-neighbour tile, [split][i].
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace g6 {

// One j-particle as given to g6_set_j_particle_ (20 doubles, 160 B).
struct JRaw {
    double x[3], v[3], a2[3], j6[3], k18[3];
    double tj, m, dtj, pad[2];
};
static_assert(sizeof(JRaw) == 160, "JRaw layout");

// One predicted j-particle (16 floats, 64 B: one scalar-cache line).  The
// ten fields the force loop reads are contiguous so they arrive in two
// scalar loads.
struct JPred {
    float xh, yh, zh;          // position, high FP32 word
    float xl, yl, zl;          // position, low FP32 word (x = xh + xl)
    float vx, vy, vz;          // velocity
    float m;                   // mass
    float pad[6];
};
static_assert(sizeof(JPred) == 64, "JPred layout");

// i-block, structure of arrays with stride `ld` (>= ni, multiple of 64).
struct IBlock {
    const float *xh, *yh, *zh, *xl, *yl, *zl, *vx, *vy, *vz;
    const int *excl;                 // address excluded for this i (-1: none)
    const double *x64, *y64, *z64;   // FP64 position (refinement, lists)
    const double *vx64, *vy64, *vz64;   // FP64 velocity (FP64 nearest tile)
    int ld;
};

// Partial results of one j-split, structure of arrays [split][ld].
struct Partials {
    double *c[7];          // ax ay az jx jy jz pot
    float *best;           // smallest FP32 r^2 (+inf: none)
    int *btile;            // first address of the j-tile holding it
    float *second;         // runner-up FP32 r^2 outside the best tile (the reduction
                           // adds the best tile's own runner-up when it rescans it)
    // The FP32 tile sums (ax ay az jx jy jz kx ky kz pot, as the force kernel
    // folded them into c[]) of the best tile: array k at tsum + k * tstride.
    // The reduction replaces the winning tile's FP32 contribution by an FP64
    // recomputation of the same pairs.
    float *tsum;
    size_t tstride;
    int ld;                // stride between splits (= IBlock ld)
};

// Final results (caller layout: double[ni][3] acc, jerk; pot[ni]; nn[ni]),
// plus flags[ni] = 1 where the nearest neighbour needed the FP64 rescan.
struct Results {
    double *acc, *jerk, *pot;
    int *nn;
    int *flags;
};

constexpr int kTile = 256;           // largest j per FP32 partial-sum tile
// Nearest-neighbour ordering by FP32 r^2 is trusted only when the runner-up
// exceeds the best by more than this relative margin (16 FP32 ulps; the
// FP32 r^2 of a double-single difference is within ~4 ulps of the exact
// value); otherwise the i is rescanned in FP64.
constexpr float kNNBand = 1.9073486328125e-06f;   // 2^-19

// Kernel launchers (g6_kernels.hip).  All enqueue on `s` and return the
// launch status.
// Scatter n j records from the pinned host mirror (device address of
// mirror) to HBM: list = n addresses, then n index fields (pinned too).
// With ibytes > 0 (a multiple of 16) the same launch copies ibytes from isrc
// (pinned host memory) to idst (HBM).
hipError_t launch_scatter(hipStream_t s, JRaw *jraw, int *jindex, const JRaw *mirror,
                          const int *list, int n, const void *isrc = nullptr,
                          void *idst = nullptr, size_t ibytes = 0);
hipError_t launch_predict(hipStream_t s, const JRaw *jraw, int nj, double ti,
                          int order, JPred *jpred, double *xp64, double *vp64,
                          int cap);
hipError_t launch_iprep_self(hipStream_t s, const JPred *jpred,
                             const double *xp64, int cap, int i_first, int ni,
                             float *ibuf, int ld);
// variant 0: scalar-load kernel (64 i per workgroup); 1: LDS-staged kernel
// (512 i per workgroup).  force_igroup() gives the i per workgroup.
// zskip: pairs whose FP32 r^2 is zero add nothing (a softening below the
// smallest normal float then counts as zero).
// wpg: waves per workgroup of the scalar-load kernel (1, 2, 4, 8 or 16),
// each on its own j-split of the same 64 i; they fold their sums in LDS, so
// force_partials() = ceil(nsplit / wpg) partial splits reach the reduction
// (the LDS kernel always writes nsplit).
// fused: predict j inside the force kernel (variant 0, at most
// kFusedMaxGroups i-groups of 64):
// the raw j records and the prediction time replace the JPred array, and
// the FP64 prediction is written to xp64 / vp64 ([3][cap]) as the predictor
// kernel writes it.
constexpr int kFusedMaxGroups = 4;
struct FusedPredict {
    const JRaw *jraw;
    double ti;
    int order;
    double *xp64, *vp64;
    int cap;
};
int force_igroup(int variant);
int force_partials(int variant, int nsplit, int wpg);
hipError_t launch_force(hipStream_t s, const JPred *jpred, int nj,
                        const IBlock &ib, int ni, float eps2, int nsplit,
                        int jchunk, int tile, const Partials &p, bool xcd_remap, int variant,
                        bool zskip, int wpg, const FusedPredict *fused);
// The reduction's nearest-neighbour rescan reads the FP64 prediction
// (xp64) and splits it into the double-single words the force kernel used.
// It also recomputes the pairs of the winning tile in FP64 (xp64, vp64,
// m64, the i-block's FP64 position and velocity, softening eps2d)
// and swaps them for that tile's FP32 sums; the tile of address t was
// summed by partial (t / jchunk) / wpg (wpg = 1 for the LDS kernel).
// zskip: as launch_force's (pairs with FP32 r^2 = 0 add nothing).
struct NearTile {
    double eps2d;
    int jchunk, wpg, zskip;
    const double *vp64;
    const double *m64;     // FP64 masses (vp64 + 3 * cap)
};
// A block's completion word: the launch that ends the block stores seq into
// word (pinned host memory) once its results are in host memory; ticket is a
// device counter, zero between launches.
struct DoneWord {
    unsigned *ticket;
    unsigned long long *word;
    unsigned long long seq;
};

hipError_t launch_reduce(hipStream_t s, const Partials &p, int nsplit, int ni,
                         int nj, const IBlock &ib, float eps2,
                         int tile, const double *xp64, int cap, const int *jindex,
                         const Results &r, int lanes, const NearTile &nt,
                         const DoneWord *done = nullptr);
// Small ABI blocks in one launch (k_small): at most kSmallMaxI i with one
// softening, FP64 pair arithmetic; each wave predicts 64 j at a time (one
// per lane) into LDS and splits its lanes into one group per i.
// The i-block and up to kSmallMaxDirty pending j updates travel in the
// kernel arguments: the lane that owns an updated address takes the record
// from there and also writes it to jraw / jindex.  Each workgroup writes its
// partial sums; the last one to finish (a ticket counter) sums them in
// workgroup order and writes the results.  write_pred: also write the FP64
// prediction (xp64, vp64, m64) as the predictor kernel does.
// i-slots of the largest instantiation; blocks of at most kSmallUseI i take
// it: per block at 9 / 12 / 16 i it measured 31.2 / 33.3 / 36.8 us against
// 35.9 / 35.8 / 36.0 us on the i-per-lane FP32 path at nj = 65536, and
// 47.2 / 48.4 / 53.0 against 52.5 / 52.3 / 52.4 us at nj = 262144
constexpr int kSmallMaxI = 16;
constexpr int kSmallUseI = 12;
constexpr int kSmallMaxDirty = 16;
constexpr int kSmallMaxWG = 256;      // workgroups per launch at most
constexpr int kSmallThreads = 256;    // threads per workgroup
constexpr int kSmallComp = 9;         // ax ay az jx jy jz pot, nearest d2, its address
// A small block as the kernel's arguments, room for MI i and MD pending
// records.  The host fills the largest form (SmallBlock); a launch copies it
// into the smallest form that holds the block, because the argument size
// shows in the launch's host time (1.1 against 3.2 KB: ~1 us per block).
template <int MI, int MD>
struct SmallArgs {
    double ti, eps2;
    int ni, nj, nd, order, write_pred, pad;
    double xi[MI][3], vi[MI][3];
    int excl[MI];                       // address excluded for each i (-1: none)
    int daddr[MD], dindex[MD];
    JRaw drec[MD];
};
using SmallBlock = SmallArgs<kSmallMaxI, kSmallMaxDirty>;
constexpr int kSmallFewDirty = 4;     // record room of the smaller launch form
static_assert(sizeof(SmallBlock) <= 3712, "kernel argument block");
// workgroups and 64-j chunks per wave for nj j-particles
void small_grid(int nj, int &wgs, int &chunks);
size_t small_part_doubles();   // partial-sum buffer of one launch
// done_flag (pinned host word, may be null): the last workgroup stores seq
// there after the results, with system-scope fences in between.
hipError_t launch_small(hipStream_t s, const SmallBlock &b, JRaw *jraw, int *jindex,
                        double *xp64, double *vp64, int cap, double *part, unsigned *ticket,
                        const Results &r, unsigned long long *done_flag = nullptr,
                        unsigned long long seq = 0);

// Potential only: pot[i] = -sum_j m_j / r_ij for the i-block (self excluded
// by address as in the force kernel); part holds nsplit * ib.ld doubles.
hipError_t launch_pot(hipStream_t s, const JPred *jpred, int nj, const IBlock &ib, int ni,
                      float eps2, int nsplit, int jchunk, int tile, double *part, double *pot,
                      bool xcd_remap);
// k-th nearest neighbours of the j-particles [i_first, i_first + ni) among
// [0, nj) (FP64 predicted positions) and their densities; 2 <= k <= 32.
// part_d / part_a hold nsplit * knn_kmax(k) * ld entries.
int knn_kmax(int k);
hipError_t launch_knn(hipStream_t s, const double *xp64, int cap, int nj, int i_first, int ni,
                      int k, int nsplit, int jchunk, double *part_d, int *part_a, int ld,
                      const JRaw *jraw, double *density, double *dk2);
// lanes per i the reduction uses by default for nsplit splits (4..64)
int reduce_lanes(int nsplit);
// Neighbour lists of the nslots i-slots listed in slots[]: counts[slot] =
// true count; the lists packed back to back in ascending slot order (at most
// nb_cap entries each), into host_packed when the packed total is at most
// host_cap entries, else into packed (device); host_info[0, nslots) = each
// list's start, host_info[nslots] = the packed total.  scratch holds
// nl_scratch_ints() ints.  slots, counts, host_info and host_packed may be
// pinned host memory (device addresses).
size_t nl_scratch_ints(int nslots, int nj);
hipError_t launch_nblist(hipStream_t s, const double *xp64, int cap, const int *jindex, int nj,
                         const IBlock &ib, const double *h2, const double *e2, const int *slots,
                         int nslots, int include_self, int nb_cap, int *counts, int *scratch,
                         int *packed, int *host_info, int *host_packed, int host_cap);

}  // namespace g6
