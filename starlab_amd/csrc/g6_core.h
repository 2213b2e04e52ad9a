// g6_core.h -- the engine shared by the GRAPE-6 front end (libg6.so,
// g6_engine.hip) and the GRAPE-4 front end (libg4x.so, g4_engine.hip).
//
// One process drives one device.  The engine keeps
//   - a pinned host mirror of j-memory (records and their index fields,
//     dirty addresses tracked) and its HBM copy,
//   - predicted j-particles in HBM (re-predicted only when ti or j changed),
//   - one i-block in flight: calc_begin() stages the i data, enqueues upload,
//     force kernels, reduction and result download on one stream and
//     returns; calc_wait() waits on an event,
//   - neighbour lists of the last i-block, built on demand (read_nl),
//   - the force-shape table of the autotuner.
// Each front end is linked into its own shared library, so each has its own
// engine state.  The front end supplies lib_tag() (message prefix) and
// env_prefix() (prefix of its environment knobs).
#pragma once

#include "g6_internal.h"

#include <string>
#include <unordered_map>
#include <vector>

namespace g6 {

const char *lib_tag();      // "libg6" / "libg4x"
const char *env_prefix();   // "G6_" / "G4_"

enum KernelId {
    K_PREDICT = 0, K_FORCE, K_REDUCE, K_REFINE, K_NBLIST, K_SCATTER, K_POT, K_KNN, K_COUNT
};

struct Options {
    int predict_order = 4;
    int npipes = 2048;
    int nb_cap = 8192;
    int waves_target = 16384;
    int nsplit = 0;
    int timing = 0;
    int device = 0;
    int xcd_remap = 1;
    int force_variant = -1;   // -1 auto, 0 scalar-load kernel, 1 LDS kernel
    int tile = 0;             // j per FP32 tile of the scalar-load kernel: 0 auto, 64/128/256
    int reduce_lanes = 0;     // lanes per i in the split reduction: 0 auto, 4..64
    int wpg = 0;              // waves per force workgroup (j-splits folded in LDS): 0 auto, 1..16
    int fused_small = 1;      // ABI blocks of one i-group predict j inside the force kernel
    int small_fp64 = 1;       // ABI blocks of at most kSmallUseI i: one FP64 launch (k_small)
    int done_word = 1;        // lasthalf polls the block's completion word, not an event
    int lds_min_ni = 1 << 30; // auto: LDS kernel from this many i on (measured
                              // slower than the scalar-load kernel at every
                              // size so far, so off unless asked for)
    // Host-side failure injection for callers' recovery paths (kira's reset
    // and retry): the next k calls of the named kind report an error; the
    // device work itself runs normally.
    int fail_lasthalf = 0;    // lasthalf returns 1 (results withheld)
    int fail_read_nl = 0;     // g6_read_neighbour_list_ returns -1
    int fail_nn = 0;          // lasthalf2 returns 0 with nnbindex[0] = -1
};

struct TimedPair { int kernel; hipEvent_t a, b; };

// Force geometry chosen per i-block size bucket: bucket b holds ni in
// (2^(b-1), 2^b].  waves = 0 marks an untuned bucket.
constexpr int kBuckets = 21;
struct Shape {
    int variant = 0, tile = 0, waves = 0, lanes = 0;
    int wpg = 0;   // waves per workgroup of the scalar-load kernel (0: built-in rule)
};

int bucket_of(int ni);

// A run of i-slots that share one softening and one self-pair rule:
// slots [off, off + n) of the staged i-block.  zskip: pairs whose FP32 r^2
// is exactly zero contribute nothing (GRAPE-4 hardware semantics; the
// GRAPE-6 front end excludes by address instead).
struct IGroup {
    int off, n;
    double eps2;   // the kernels use it rounded to FP32, the FP64 nearest tile as is
    bool zskip;
};

struct Engine {
    Options opt;
    bool env_read = false;
    bool open = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    // completion word of an ABI block: the workgroup that finishes the
    // block's last launch (k_small, or the last k_reduce) stores the block's
    // sequence number here (pinned, coherent) after the results, so lasthalf
    // can poll it instead of waiting for ev_done's completion signal
    unsigned long long *h_flag = nullptr;
    unsigned long long *h_flag_dev = nullptr;
    unsigned long long flag_seq = 0;
    bool flag_pending = false;
    std::string last_error;

    // ---- j-memory ---------------------------------------------------------
    int hcap = 0;               // host mirror capacity (addresses)
    JRaw *h_jraw = nullptr;     // pinned
    const JRaw *h_jraw_dev = nullptr;   // its device address (read by k_scatter)
    int *h_jindex = nullptr;    // pinned
    int n_addr = 0;             // 1 + highest address ever set
    std::vector<int> dirty;
    std::vector<uint8_t> dirty_flag;
    // index field -> address of the j that holds it (self-exclusion): a flat
    // array for indices in [0, kDenseIndex), a hash map for any other
    std::vector<int> addr_by_index;
    std::unordered_map<int, int> addr_of_index;

    int dcap = 0;               // device capacity
    JRaw *d_jraw = nullptr;
    JPred *d_jpred = nullptr;
    double *d_xp64 = nullptr;   // [3][dcap]
    double *d_vp64 = nullptr;   // [3][dcap] velocities, then the FP64 masses m64[dcap]
    int *d_jindex = nullptr;

    int scap = 0;               // scatter staging capacity (records)
    char *h_stage = nullptr;    // pinned: n records, n addresses, n index fields;
    const char *h_stage_dev = nullptr;   // its device address (k_scatter reads it)

    bool pred_valid = false;   // JPred, xp64, vp64 hold the prediction at pred_ti
    double pred_ti = 0;
    int pred_nj = 0, pred_order = 0;
    bool xp_valid = false;     // xp64, vp64 hold the prediction at xp_ti (the predictor
    double xp_ti = 0;          // or a fused small block wrote them last)
    int xp_nj = 0;
    double ti = 0;

    // ---- i-block ----------------------------------------------------------
    int icap = 0;               // slots (multiple of 64)
    char *h_ibuf = nullptr;     // pinned, layout of ibuf_bytes()
    const char *h_ibuf_dev = nullptr;   // its device address (read by k_scatter)
    char *d_ibuf = nullptr;
    int *h_slots = nullptr;     // pinned: i-slots with neighbour lists

    // ---- partial sums -----------------------------------------------------
    size_t pbytes = 0;
    char *d_part = nullptr;
    // k_small: its workgroup partials (small_part_doubles()), then its
    // ticket counter (zero between launches)
    double *d_small = nullptr;
    unsigned *d_done_ticket = nullptr;   // k_reduce's completion ticket (zero between launches)

    // ---- results ----------------------------------------------------------
    int rcap = 0;
    char *d_res = nullptr;      // acc[3n] jerk[3n] pot[n] nn[n] flags[n]
    char *h_res = nullptr;      // pinned; the ABI path's reduction writes it directly
    char *h_res_dev = nullptr;  // its device address

    // Pinned buffers used by queued device work: h_stage (the scatter list),
    // the mirror h_jraw / h_jindex (records read by k_scatter or a full
    // upload), h_ibuf (read by k_scatter), h_res (written by the reduction).
    // Each such enqueue takes the next op_seq; the buffer remembers it.
    // synced_seq is the last op_seq known complete (lasthalf's wait on
    // ev_done, which follows done_seq).  Host code that rewrites or frees a
    // buffer waits only if that buffer's op is newer than synced_seq.
    unsigned long long op_seq = 0, synced_seq = 0, done_seq = 0;
    unsigned long long stage_seq = 0, ibuf_seq = 0, res_seq = 0;
    int deferred_scatter = 0;   // records listed in h_stage, not yet launched
    unsigned long long mirror_seq = 0;   // last queued op that reads h_jraw / h_jindex

    // ---- neighbour lists --------------------------------------------------
    size_t nbcap_entries = 0;
    int nbcount_cap = 0;
    int *d_nblist = nullptr;     // lists too long for the pinned buffer
    int *d_nlscratch = nullptr;  // per (slot, split) counts, list bases, packed total
    size_t nlscratch_cap = 0;
    // pinned buffers the list kernels use directly (device addresses *_dev)
    int *h_slots_dev = nullptr, *h_nbcount_dev = nullptr, *h_nblist_dev = nullptr;
    int *h_nlinfo = nullptr, *h_nlinfo_dev = nullptr;   // list starts + packed total
    int nlinfo_cap = 0;
    int *h_nbcount = nullptr, *h_nblist = nullptr;   // pinned
    size_t h_nblist_cap = 0;
    std::vector<size_t> nb_offset;
    bool nb_valid = false;

    // ---- state of the last calculation -------------------------------------
    bool calc_pending = false;
    bool calc_failed = false;
    bool sweep_results = false;   // d_res holds the last calculation (g6x_self_sweep)
    int last_ni = 0, last_nj = 0;
    int last_refinements = 0;
    std::vector<double> last_h2;
    int last_include_self = 1;
    int geom_wg = 0, geom_nsplit = 0, geom_jchunk = 0, geom_variant = 0, geom_tile = 0;
    int geom_lanes = 0, geom_wpg = 1;
    Shape tuned[kBuckets];
    bool forcing_shape = false;   // the autotuner's candidate overrides everything
    Shape forced;

    // ---- timing -------------------------------------------------------------
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimedPair> pending;
    double k_ms[K_COUNT] = {0};
    int k_calls[K_COUNT] = {0};
};

extern Engine E;

int fail(int code, const char *fmt, ...);

#define HIPCHK(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::g6::fail(-1, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

void read_env();
void harvest_timing();

// j-memory
int set_j(int address, int index, double tj, double dtj, double mass,
          const double *k18, const double *j6, const double *a2,
          const double *v, const double *x);
int excl_address(int index);
// defer: leave a packed scatter in h_stage for calc_begin to launch together
// with the i-block upload (deferred_scatter records); any other flush_j
// launches a deferred scatter first.
// noalloc: fail instead of growing any buffer (a dirty list longer than the
// staging buffer then goes up as one copy of the used range)
int flush_j(int nj, bool defer = false, bool noalloc = false);
int predict(int nj, bool force);

// i-block and results (host or device copies share one layout)
int ensure_i(int ni);
size_t results_bytes(int ni);
IBlock iblock_at(char *base, int ld);
Results results_at(char *base, int ni);
int count_refined(const Results &h, int ni);
// fused (optional): predict j inside the force kernel instead of reading the
// JPred array (one i-group; the caller has not run the predictor)
// part (optional): device memory for the j-split partial sums, at least
// partials_bytes(ib.ld, partial splits of the geometry); nullptr = the
// library's own (grown) buffer.
int enqueue_force(int ni, int nj, double eps2, const IBlock &ib, const Results &r,
                  bool zskip = false, const FusedPredict *fused = nullptr, char *part = nullptr,
                  const DoneWord *done = nullptr);

// One i-block through the force path.  Slot k has position x[k], velocity
// v[k], excluded j address excl[k] (-1: none; nullptr: none for all), and
// for the neighbour list the squared radius h2[k] and the offset e2[k]
// added to d^2 before the comparison (nullptr: 0).  groups partition the
// slots by softening; include_self decides whether a list keeps j ==
// excl[k].  Asynchronous: results are read with calc_wait().
int calc_begin(const char *who, int nj, int ni, const double (*x)[3], const double (*v)[3],
               const int *excl, const double *h2, const double *e2, const IGroup *groups,
               int ngroups, int include_self);
// Wait for the i-block begun last; returns its host results (nullptr and an
// error message on failure).
const Results *calc_wait(const char *who);

// Neighbour lists of the last i-block: read_nl builds them for every slot
// with h2 > 0 and returns 1 if some list exceeded the store (nb_cap
// entries), 0 if not, < 0 on error.  nl_count / nl_list give one slot's
// true count and its stored entries (min(count, nb_cap)).
int read_nl();
int nl_count(int slot);
const int *nl_list(int slot);
int nl_max_count();
int get_nl(int pipe, int max_length, int *nblen, int *nbl);

int ensure_open(const char *who);
int do_open(int dev);
int do_close();
void reset_state();

// device-resident sweeps and the autotuner
int self_block(int nj, int i_first, int ni, IBlock &ib);
// g6x_self_sweep without allocation: scratch (i-block, flags, partials) in
// the caller's workspace ws (ws_bytes, device memory), results to the device
// arrays of out (acc, jerk, pot, nn; its flags are ignored).  Fails, with a
// message, when ws_bytes < sweep_workspace_bytes(nj, ni), when j-memory does
// not cover nj addresses or is not resident on the device.  Asynchronous on
// the library stream.
size_t sweep_workspace_bytes(int nj, int ni);
int self_sweep_into(const char *who, int nj, int i_first, int ni, double ti, double eps2,
                    const Results &out, char *ws, size_t ws_bytes);
// pot of the j-particles [i_first, i_first + ni) against [0, nj) at time ti
// (self excluded), to the host array pot[ni]
int potential_sweep(int nj, int i_first, int ni, double ti, double eps2, double *pot);
// k-th nearest-neighbour distance^2 and density (compute_density) of the
// j-particles [i_first, i_first + ni) among [0, nj) at time ti
int density_sweep(int nj, int i_first, int ni, double ti, int k, double *density, double *dk2);
int autotune(int nj, int ni_max, double eps2);
int save_tuning(const char *path);
int load_tuning(const char *path);

}  // namespace g6
