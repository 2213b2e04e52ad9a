// g6_engine.hip -- host side of libg6 for MI355X: the GRAPE-6 C ABI
// (include/grape6.h) and its extensions (include/g6x.h).
//
// One process drives one device.  The library keeps
//   - a pinned host mirror of j-memory (records written by
//     g6_set_j_particle_, dirty addresses tracked) and its HBM copy,
//   - predicted j-particles in HBM (re-predicted only when ti or j changed),
//   - one i-block in flight: g6calc_firsthalf_ stages the i data, enqueues
//     upload + kernels + result download on one stream and returns;
//     g6calc_lasthalf(2)_ waits on an event and unpacks.
// The call sequence kira uses (hdyn_grape6.C:474-910, 1991-2181, 3067-3411)
// is open -> set_j* -> set_ti -> firsthalf -> lasthalf2 -> [read_nl ->
// get_nl*], repeated, with close/re-open and resets at any point.
#include "g6_internal.h"
#include "../../include/grape6.h"
#include "../../include/g6x.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace g6 {
namespace {

enum KernelId { K_PREDICT = 0, K_FORCE, K_REDUCE, K_REFINE, K_NBLIST, K_SCATTER, K_COUNT };

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
    int lds_min_ni = 1 << 30; // auto: LDS kernel from this many i on (measured
                              // slower than the scalar-load kernel at every
                              // size so far, so off unless asked for)
};

struct TimedPair { int kernel; hipEvent_t a, b; };

// Force geometry chosen per i-block size bucket: bucket b holds ni in
// (2^(b-1), 2^b].  waves = 0 marks an untuned bucket.
constexpr int kBuckets = 21;
struct Shape {
    int variant = 0, tile = 0, waves = 0, lanes = 0;
};

int bucket_of(int ni)
{
    int b = 0;
    while (b < kBuckets - 1 && (1 << b) < ni) ++b;
    return b;
}

struct Engine {
    Options opt;
    bool env_read = false;
    bool open = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    std::string last_error;

    // ---- j-memory ---------------------------------------------------------
    int hcap = 0;               // host mirror capacity (addresses)
    JRaw *h_jraw = nullptr;     // pinned
    int *h_jindex = nullptr;    // pinned
    int n_addr = 0;             // 1 + highest address ever set
    std::vector<int> dirty;
    std::vector<uint8_t> dirty_flag;
    std::unordered_map<int, int> addr_of_index;

    int dcap = 0;               // device capacity
    JRaw *d_jraw = nullptr;
    JPred *d_jpred = nullptr;
    double *d_xp64 = nullptr;   // [3][dcap]
    double *d_vp64 = nullptr;   // [3][dcap]
    int *d_jindex = nullptr;

    int scap = 0;               // scatter staging capacity (records)
    char *h_stage = nullptr;    // pinned: n records, n addresses, n index fields
    char *d_stage = nullptr;

    bool pred_valid = false;
    double pred_ti = 0;
    int pred_nj = 0, pred_order = 0;
    double ti = 0;

    // ---- i-block ----------------------------------------------------------
    int icap = 0;               // slots (multiple of 64)
    char *h_ibuf = nullptr;     // pinned: 10*icap*4 + 4*icap*8 bytes
    char *d_ibuf = nullptr;
    int *d_slots = nullptr; int *h_slots = nullptr;

    // ---- partial sums -----------------------------------------------------
    size_t pbytes = 0;
    char *d_part = nullptr;

    // ---- results ----------------------------------------------------------
    int rcap = 0;
    char *d_res = nullptr;      // acc[3n] jerk[3n] pot[n] nn[n] flags[n]
    char *h_res = nullptr;      // pinned

    // ---- neighbour lists --------------------------------------------------
    size_t nbcap_entries = 0;
    int *d_nbcount = nullptr, *d_nblist = nullptr;
    int *h_nbcount = nullptr, *h_nblist = nullptr;
    size_t h_nblist_cap = 0;
    std::vector<size_t> nb_offset;
    bool nb_valid = false;

    // ---- state of the last calculation -------------------------------------
    bool calc_pending = false;
    bool calc_failed = false;
    int last_ni = 0, last_nj = 0;
    int last_refinements = 0;
    std::vector<double> last_h2;
    int last_include_self = 1;
    int geom_wg = 0, geom_nsplit = 0, geom_jchunk = 0, geom_variant = 0, geom_tile = 0;
    int geom_lanes = 0;
    Shape tuned[kBuckets];
    bool forcing_shape = false;   // the autotuner's candidate overrides everything
    Shape forced;

    // ---- timing -------------------------------------------------------------
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimedPair> pending;
    double k_ms[K_COUNT] = {0};
    int k_calls[K_COUNT] = {0};
};

Engine E;

// ---------------------------------------------------------------------------
int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    E.last_error = buf;
    fprintf(stderr, "libg6: %s\n", buf);
    return code;
}

#define HIPCHK(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return fail(-1, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

long long env_ll(const char *name, long long dflt)
{
    const char *s = getenv(name);
    if (!s || !*s) return dflt;
    return atoll(s);
}

void read_env()
{
    if (E.env_read) return;
    E.env_read = true;
    Options &o = E.opt;
    o.predict_order = (int)env_ll("G6_PREDICT_ORDER", o.predict_order);
    o.npipes = (int)env_ll("G6_NPIPES", o.npipes);
    o.nb_cap = (int)env_ll("G6_NB_CAP", o.nb_cap);
    o.waves_target = (int)env_ll("G6_WAVES", o.waves_target);
    o.nsplit = (int)env_ll("G6_NSPLIT", o.nsplit);
    o.timing = (int)env_ll("G6_TIMING", o.timing);
    o.device = (int)env_ll("G6_DEVICE", o.device);
    o.xcd_remap = (int)env_ll("G6_XCD_REMAP", o.xcd_remap);
    o.force_variant = (int)env_ll("G6_FORCE_VARIANT", o.force_variant);
    o.lds_min_ni = (int)env_ll("G6_LDS_MIN_NI", o.lds_min_ni);
    o.tile = (int)env_ll("G6_TILE", o.tile);
    o.reduce_lanes = (int)env_ll("G6_REDUCE_LANES", o.reduce_lanes);
    if (o.npipes < 1) o.npipes = 1;
    if (o.nb_cap < 1) o.nb_cap = 1;
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
hipEvent_t ev_get()
{
    if (!E.ev_pool.empty()) {
        hipEvent_t e = E.ev_pool.back();
        E.ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct KTimer {
    int k;
    hipEvent_t a = nullptr;
    explicit KTimer(int kernel) : k(kernel)
    {
        if (E.opt.timing && (a = ev_get())) hipEventRecord(a, E.stream);
    }
    void stop()
    {
        if (!a) return;
        hipEvent_t b = ev_get();
        if (!b) return;
        hipEventRecord(b, E.stream);
        E.pending.push_back({k, a, b});
        a = nullptr;
    }
};

void harvest_timing()
{
    for (auto &p : E.pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            E.k_ms[p.kernel] += ms;
            E.k_calls[p.kernel] += 1;
        }
        E.ev_pool.push_back(p.a);
        E.ev_pool.push_back(p.b);
    }
    E.pending.clear();
}

// ---------------------------------------------------------------------------
// memory management
// ---------------------------------------------------------------------------
template <class T> void dfree(T *&p) { if (p) { (void)hipFree(p); p = nullptr; } }
template <class T> void hfree(T *&p) { if (p) { (void)hipHostFree(p); p = nullptr; } }

int grow_host_mirror(int need)
{
    if (need <= E.hcap) return 0;
    int cap = std::max(need, std::max(1024, E.hcap * 2));
    JRaw *nr = nullptr;
    int *ni = nullptr;
    HIPCHK(hipHostMalloc((void **)&nr, sizeof(JRaw) * (size_t)cap, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&ni, sizeof(int) * (size_t)cap, hipHostMallocDefault));
    memset(nr, 0, sizeof(JRaw) * (size_t)cap);
    for (int k = 0; k < cap; ++k) ni[k] = -1;
    if (E.hcap) {
        memcpy(nr, E.h_jraw, sizeof(JRaw) * (size_t)E.hcap);
        memcpy(ni, E.h_jindex, sizeof(int) * (size_t)E.hcap);
    }
    hfree(E.h_jraw);
    hfree(E.h_jindex);
    E.h_jraw = nr;
    E.h_jindex = ni;
    E.hcap = cap;
    E.dirty_flag.resize(cap, 0);
    return 0;
}

void free_device_j()
{
    dfree(E.d_jraw); dfree(E.d_jpred); dfree(E.d_xp64); dfree(E.d_vp64); dfree(E.d_jindex);
    E.dcap = 0;
    E.pred_valid = false;
}

// Device j capacity >= need; a (re)allocation uploads the whole mirror.
int ensure_device_j(int need)
{
    if (need <= E.dcap) return 0;
    int cap = std::max(need, std::max(1024, E.dcap * 2));
    cap = (cap + 255) & ~255;
    free_device_j();
    HIPCHK(hipMalloc((void **)&E.d_jraw, sizeof(JRaw) * (size_t)cap));
    HIPCHK(hipMalloc((void **)&E.d_jpred, sizeof(JPred) * (size_t)cap));
    HIPCHK(hipMalloc((void **)&E.d_xp64, sizeof(double) * 3 * (size_t)cap));
    HIPCHK(hipMalloc((void **)&E.d_vp64, sizeof(double) * 3 * (size_t)cap));
    HIPCHK(hipMalloc((void **)&E.d_jindex, sizeof(int) * (size_t)cap));
    HIPCHK(hipMemsetAsync(E.d_jraw, 0, sizeof(JRaw) * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_jpred, 0, sizeof(JPred) * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_xp64, 0, sizeof(double) * 3 * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_vp64, 0, sizeof(double) * 3 * (size_t)cap, E.stream));
    E.dcap = cap;
    // everything set so far must reach the new arrays
    for (int a : E.dirty) E.dirty_flag[a] = 0;
    E.dirty.clear();
    const int n = std::min(E.n_addr, E.hcap);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(E.d_jraw, E.h_jraw, sizeof(JRaw) * (size_t)n,
                              hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(E.d_jindex, E.h_jindex, sizeof(int) * (size_t)n,
                              hipMemcpyHostToDevice, E.stream));
    }
    return 0;
}

int ensure_stage(int n)
{
    if (n <= E.scap) return 0;
    int cap = std::max(n, std::max(256, E.scap * 2));
    hfree(E.h_stage);
    dfree(E.d_stage);
    E.scap = 0;
    const size_t bytes = (sizeof(JRaw) + 2 * sizeof(int)) * (size_t)cap;
    HIPCHK(hipHostMalloc((void **)&E.h_stage, bytes, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E.d_stage, bytes));
    E.scap = cap;
    return 0;
}

// Push dirty j records to HBM: the whole used range when many changed,
// otherwise a packed list scattered by a kernel.
int flush_j(int nj)
{
    if (nj > E.n_addr) {
        // addresses never set are zero-mass particles at the origin
        if (grow_host_mirror(nj)) return -1;
        E.n_addr = nj;
    }
    if (ensure_device_j(E.n_addr)) return -1;
    const int nd = (int)E.dirty.size();
    if (nd == 0) return 0;
    if ((size_t)nd * 4 >= (size_t)E.n_addr) {
        HIPCHK(hipMemcpyAsync(E.d_jraw, E.h_jraw, sizeof(JRaw) * (size_t)E.n_addr,
                              hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(E.d_jindex, E.h_jindex, sizeof(int) * (size_t)E.n_addr,
                              hipMemcpyHostToDevice, E.stream));
    } else {
        if (ensure_stage(nd)) return -1;
        // the staging buffers may still be read by an earlier copy
        HIPCHK(hipStreamSynchronize(E.stream));
        // one packed upload: records, then addresses, then index fields
        JRaw *rec = reinterpret_cast<JRaw *>(E.h_stage);
        int *addr = reinterpret_cast<int *>(rec + nd);
        for (int k = 0; k < nd; ++k) {
            const int a = E.dirty[k];
            rec[k] = E.h_jraw[a];
            addr[k] = a;
            addr[nd + k] = E.h_jindex[a];
        }
        HIPCHK(hipMemcpyAsync(E.d_stage, E.h_stage, (sizeof(JRaw) + 2 * sizeof(int)) * (size_t)nd,
                              hipMemcpyHostToDevice, E.stream));
        KTimer t(K_SCATTER);
        HIPCHK(launch_scatter(E.stream, E.d_jraw, E.d_jindex,
                              reinterpret_cast<const JRaw *>(E.d_stage), nd));
        t.stop();
    }
    for (int a : E.dirty) E.dirty_flag[a] = 0;
    E.dirty.clear();
    E.pred_valid = false;
    return 0;
}

int predict(int nj, bool force)
{
    if (!force && E.pred_valid && E.pred_ti == E.ti && E.pred_nj >= nj &&
        E.pred_order == E.opt.predict_order)
        return 0;
    KTimer t(K_PREDICT);
    HIPCHK(launch_predict(E.stream, E.d_jraw, nj, E.ti, E.opt.predict_order,
                          E.d_jpred, E.d_xp64, E.d_vp64, E.dcap));
    t.stop();
    E.pred_valid = true;
    E.pred_ti = E.ti;
    E.pred_nj = nj;
    E.pred_order = E.opt.predict_order;
    return 0;
}

// i-block buffer: 10 four-byte arrays then 4 double arrays (x64 y64 z64 h2).
size_t ibuf_bytes(int cap) { return (size_t)cap * (10 * 4 + 4 * 8); }

// results: 7 doubles, nn and the refinement flag per i
size_t results_bytes(int ni) { return (size_t)ni * (7 * sizeof(double) + 2 * sizeof(int)); }

int ensure_i(int ni)
{
    if (ni <= E.icap) return 0;
    int cap = std::max((ni + 63) & ~63, 256);
    hfree(E.h_ibuf); dfree(E.d_ibuf); hfree(E.h_slots); dfree(E.d_slots);
    hfree(E.h_res); dfree(E.d_res);
    E.icap = 0;
    HIPCHK(hipHostMalloc((void **)&E.h_ibuf, ibuf_bytes(cap), hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E.d_ibuf, ibuf_bytes(cap)));
    HIPCHK(hipHostMalloc((void **)&E.h_slots, sizeof(int) * (size_t)cap, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E.d_slots, sizeof(int) * (size_t)cap));
    const size_t rb = results_bytes(cap);
    HIPCHK(hipHostMalloc((void **)&E.h_res, rb, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E.d_res, rb));
    E.icap = cap;
    return 0;
}

IBlock iblock_at(char *base, int ld)
{
    IBlock b;
    const float *f = reinterpret_cast<const float *>(base);
    b.xh = f + 0 * (size_t)ld; b.yh = f + 1 * (size_t)ld; b.zh = f + 2 * (size_t)ld;
    b.xl = f + 3 * (size_t)ld; b.yl = f + 4 * (size_t)ld; b.zl = f + 5 * (size_t)ld;
    b.vx = f + 6 * (size_t)ld; b.vy = f + 7 * (size_t)ld; b.vz = f + 8 * (size_t)ld;
    b.excl = reinterpret_cast<const int *>(f + 9 * (size_t)ld);
    const double *d = reinterpret_cast<const double *>(base + (size_t)ld * 40);
    b.x64 = d; b.y64 = d + ld; b.z64 = d + 2 * (size_t)ld;
    b.ld = ld;
    return b;
}

double *h2_at(char *base, int ld)
{
    return reinterpret_cast<double *>(base + (size_t)ld * 40) + 3 * (size_t)ld;
}

Results results_at(char *base, int ni)
{
    Results r;
    double *d = reinterpret_cast<double *>(base);
    r.acc = d;
    r.jerk = d + 3 * (size_t)ni;
    r.pot = d + 6 * (size_t)ni;
    r.nn = reinterpret_cast<int *>(d + 7 * (size_t)ni);
    r.flags = r.nn + ni;
    return r;
}

int count_refined(const Results &h, int ni)
{
    int c = 0;
    for (int k = 0; k < ni; ++k) c += h.flags[k];
    return c;
}

int ensure_partials(size_t bytes)
{
    if (bytes <= E.pbytes) return 0;
    dfree(E.d_part);
    E.pbytes = 0;
    size_t b = std::max(bytes, E.pbytes * 2);
    HIPCHK(hipMalloc((void **)&E.d_part, b));
    E.pbytes = b;
    return 0;
}

Partials partials_at(int ld, int nsplit)
{
    Partials p;
    const size_t n = (size_t)ld * nsplit;
    double *d = reinterpret_cast<double *>(E.d_part);
    for (int c = 0; c < 7; ++c) p.c[c] = d + c * n;
    float *u = reinterpret_cast<float *>(d + 7 * n);
    p.best = u;
    p.btile = reinterpret_cast<int *>(u + n);
    p.second = u + 2 * n;
    p.ld = ld;
    return p;
}

size_t partials_bytes(int ld, int nsplit) { return (size_t)ld * nsplit * (7 * 8 + 3 * 4); }

struct Geometry {
    int variant, tile, nsplit, jchunk, lanes;
};

// Force geometry: kernel variant, j-tile length (FP32 partial-sum span),
// j-split and the reduction's lanes per i.  Precedence: the autotuner's
// candidate while it measures, explicit options, the tuned table for this
// i-block size, then the built-in rule (enough waves to fill the chip; short
// tiles when a small i-block must spread its j over many waves).
Geometry choose_geometry(int ni, int nj)
{
    Shape sh;
    const Shape &t = E.tuned[bucket_of(ni)];
    if (E.forcing_shape) {
        sh = E.forced;
    } else {
        if (t.waves > 0) sh = t;
        if (E.opt.force_variant == 0 || E.opt.force_variant == 1) sh.variant = E.opt.force_variant;
        else if (t.waves == 0) sh.variant = ni >= E.opt.lds_min_ni ? 1 : 0;
        if (E.opt.tile == 64 || E.opt.tile == 128 || E.opt.tile == kTile) sh.tile = E.opt.tile;
        if (E.opt.reduce_lanes > 0) sh.lanes = E.opt.reduce_lanes;
        if (t.waves == 0 || E.opt.waves_target != Options().waves_target)
            sh.waves = E.opt.waves_target;
    }
    Geometry g;
    g.variant = sh.variant;
    const int per_wave = force_igroup(g.variant) / (g.variant == 1 ? 4 : 1);
    const int n_iw = std::max(1, (ni + per_wave - 1) / per_wave);
    const int target = std::max(1, sh.waves);
    if (g.variant == 1) g.tile = kTile;
    else if (sh.tile == 64 || sh.tile == 128 || sh.tile == kTile) g.tile = sh.tile;
    else g.tile = (long long)n_iw * ((nj + kTile - 1) / kTile) >= target ? kTile : 64;
    const int ntiles = std::max(1, (nj + g.tile - 1) / g.tile);
    int sp = (!E.forcing_shape && E.opt.nsplit > 0) ? E.opt.nsplit
                                                    : std::max(1, (target + n_iw - 1) / n_iw);
    sp = std::min(sp, ntiles);
    const int tiles_per = (ntiles + sp - 1) / sp;
    g.jchunk = tiles_per * g.tile;
    g.nsplit = std::max(1, (nj + g.jchunk - 1) / g.jchunk);
    g.lanes = sh.lanes > 0 ? sh.lanes : reduce_lanes(g.nsplit);
    return g;
}

// Enqueue force + reduction (with the nearest-neighbour refinement) for an
// i-block already in d_ibuf.
int enqueue_force(int ni, int nj, float eps2, const IBlock &ib, const Results &r)
{
    const Geometry g = choose_geometry(ni, nj);
    const int ld = ib.ld;
    if (ensure_partials(partials_bytes(ld, g.nsplit))) return -1;
    Partials p = partials_at(ld, g.nsplit);
    E.geom_nsplit = g.nsplit;
    E.geom_jchunk = g.jchunk;
    E.geom_tile = g.tile;
    E.geom_lanes = g.lanes;
    E.geom_wg = ((ni + force_igroup(g.variant) - 1) / force_igroup(g.variant)) * g.nsplit;
    E.geom_variant = g.variant;
    {
        KTimer t(K_FORCE);
        HIPCHK(launch_force(E.stream, E.d_jpred, nj, ib, ni, eps2, g.nsplit, g.jchunk, g.tile,
                            p, E.opt.xcd_remap != 0, g.variant));
        t.stop();
    }
    {
        KTimer t(K_REDUCE);
        HIPCHK(launch_reduce(E.stream, p, g.nsplit, ni, E.d_jpred, nj, ib, eps2, g.tile,
                             E.d_xp64, E.dcap, E.d_jindex, r, g.lanes));
        t.stop();
    }
    return 0;
}

int ensure_open(const char *who)
{
    if (!E.open) return fail(-1, "%s: device not open (call g6_open_ first)", who);
    return 0;
}

int load_tuning(const char *path);

int do_open(int dev)
{
    read_env();
    if (E.open) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(-1, "g6_open: no HIP device available (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "0 devices");
    if (dev < 0 || dev >= n) dev = 0;
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreate(&E.stream));
    if (const char *tf = getenv("G6_TUNE_FILE")) {
        if (*tf && load_tuning(tf) < 0) E.last_error.clear();   // a missing file is not an error
    }
    HIPCHK(hipEventCreateWithFlags(&E.ev_done, hipEventDisableTiming));
    E.device = dev;
    E.open = true;
    E.pred_valid = false;
    E.calc_pending = false;
    E.nb_valid = false;
    // j records set before this open (or before a close) are re-sent
    E.dcap = 0;
    E.dirty.clear();
    std::fill(E.dirty_flag.begin(), E.dirty_flag.end(), 0);
    return 0;
}

int do_close()
{
    if (!E.open) return 0;
    (void)hipStreamSynchronize(E.stream);
    harvest_timing();
    free_device_j();
    dfree(E.d_stage);
    hfree(E.h_stage);
    E.scap = 0;
    hfree(E.h_ibuf); dfree(E.d_ibuf); hfree(E.h_slots); dfree(E.d_slots);
    hfree(E.h_res); dfree(E.d_res);
    E.icap = 0;
    dfree(E.d_part); E.pbytes = 0;
    dfree(E.d_nbcount); dfree(E.d_nblist); hfree(E.h_nbcount); hfree(E.h_nblist);
    E.nbcap_entries = 0; E.h_nblist_cap = 0;
    for (hipEvent_t ev : E.ev_pool) (void)hipEventDestroy(ev);
    E.ev_pool.clear();
    if (E.ev_done) (void)hipEventDestroy(E.ev_done);
    E.ev_done = nullptr;
    if (E.stream) (void)hipStreamDestroy(E.stream);
    E.stream = nullptr;
    E.open = false;
    E.calc_pending = false;
    E.nb_valid = false;
    return 0;
}

int set_j(int address, int index, double tj, double dtj, double mass,
          const double *k18, const double *j6, const double *a2,
          const double *v, const double *x)
{
    if (address < 0) return fail(-1, "g6_set_j_particle: negative address %d", address);
    if (grow_host_mirror(address + 1)) return -1;
    JRaw &r = E.h_jraw[address];
    for (int k = 0; k < 3; ++k) {
        r.x[k] = x[k]; r.v[k] = v[k]; r.a2[k] = a2[k]; r.j6[k] = j6[k]; r.k18[k] = k18[k];
    }
    r.tj = tj;
    r.m = mass;
    r.dtj = dtj;
    r.pad[0] = r.pad[1] = 0;
    const int old = E.h_jindex[address];
    if (old != index) {
        auto it = E.addr_of_index.find(old);
        if (it != E.addr_of_index.end() && it->second == address) E.addr_of_index.erase(it);
        E.h_jindex[address] = index;
    }
    E.addr_of_index[index] = address;
    if (address + 1 > E.n_addr) E.n_addr = address + 1;
    if (!E.dirty_flag[address]) {
        E.dirty_flag[address] = 1;
        E.dirty.push_back(address);
    }
    return 0;
}

int excl_address(int index)
{
    auto it = E.addr_of_index.find(index);
    return it == E.addr_of_index.end() ? -1 : it->second;
}

int firsthalf(int nj, int ni, const int *index, const double (*xi)[3],
              const double (*vi)[3], double eps2, const double *h2)
{
    E.calc_pending = false;
    E.calc_failed = true;
    E.nb_valid = false;
    if (ensure_open("g6calc_firsthalf")) return -1;
    if (ni <= 0 || ni > E.opt.npipes)
        return fail(-1, "g6calc_firsthalf: ni = %d outside [1, npipes = %d]", ni, E.opt.npipes);
    if (nj <= 0) return fail(-1, "g6calc_firsthalf: nj = %d", nj);
    if (flush_j(nj)) return -1;
    if (predict(nj, false)) return -1;
    if (ensure_i(ni)) return -1;
    const int ld = (ni + 63) & ~63;
    // The pinned i buffer may still be read by the previous upload.
    HIPCHK(hipStreamSynchronize(E.stream));
    float *f = reinterpret_cast<float *>(E.h_ibuf);
    int *ex = reinterpret_cast<int *>(f + 9 * (size_t)ld);
    double *d = reinterpret_cast<double *>(E.h_ibuf + (size_t)ld * 40);
    double *hh = d + 3 * (size_t)ld;
    E.last_h2.assign(h2, h2 + ni);
    for (int k = 0; k < ld; ++k) {
        const int s = k < ni ? k : ni - 1;
        for (int c = 0; c < 3; ++c) {
            const double x = xi[s][c];
            const float hi = (float)x;
            f[(0 + c) * (size_t)ld + k] = hi;
            f[(3 + c) * (size_t)ld + k] = (float)(x - (double)hi);
            f[(6 + c) * (size_t)ld + k] = (float)vi[s][c];
            d[c * (size_t)ld + k] = x;
        }
        ex[k] = k < ni ? excl_address(index[k]) : -1;
        hh[k] = k < ni ? h2[k] : 0.0;
    }
    HIPCHK(hipMemcpyAsync(E.d_ibuf, E.h_ibuf, ibuf_bytes(ld), hipMemcpyHostToDevice, E.stream));
    IBlock ib = iblock_at(E.d_ibuf, ld);
    Results r = results_at(E.d_res, ni);
    if (enqueue_force(ni, nj, (float)eps2, ib, r)) return -1;
    HIPCHK(hipMemcpyAsync(E.h_res, E.d_res, results_bytes(ni), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipEventRecord(E.ev_done, E.stream));
    E.last_ni = ni;
    E.last_nj = nj;
    E.last_include_self = E.opt.npipes > 48;
    E.calc_pending = true;
    E.calc_failed = false;
    return 0;
}

int lasthalf(int ni, double (*acc)[3], double (*jerk)[3], double *pot, int *nn)
{
    if (!E.calc_pending) {
        if (E.calc_failed) return fail(1, "g6calc_lasthalf: the first half failed");
        return fail(1, "g6calc_lasthalf: no calculation in flight");
    }
    E.calc_pending = false;
    HIPCHK(hipEventSynchronize(E.ev_done));
    HIPCHK(hipGetLastError());
    harvest_timing();
    const int n = std::min(ni, E.last_ni);
    Results h = results_at(E.h_res, E.last_ni);
    E.last_refinements = count_refined(h, E.last_ni);
    memcpy(acc, h.acc, sizeof(double) * 3 * (size_t)n);
    memcpy(jerk, h.jerk, sizeof(double) * 3 * (size_t)n);
    memcpy(pot, h.pot, sizeof(double) * (size_t)n);
    if (nn) memcpy(nn, h.nn, sizeof(int) * (size_t)n);
    return 0;
}

int ensure_nb(int ni)
{
    const size_t need = (size_t)ni * (size_t)E.opt.nb_cap;
    if (need > E.nbcap_entries) {
        dfree(E.d_nblist); dfree(E.d_nbcount); hfree(E.h_nbcount);
        E.nbcap_entries = 0;
        HIPCHK(hipMalloc((void **)&E.d_nblist, sizeof(int) * need));
        HIPCHK(hipMalloc((void **)&E.d_nbcount, sizeof(int) * (size_t)std::max(ni, 1)));
        HIPCHK(hipHostMalloc((void **)&E.h_nbcount, sizeof(int) * (size_t)std::max(ni, 1),
                             hipHostMallocDefault));
        E.nbcap_entries = need;
    }
    return 0;
}

int read_nl()
{
    if (ensure_open("g6_read_neighbour_list")) return -1;
    if (E.last_ni <= 0 || E.calc_failed) return fail(-1, "g6_read_neighbour_list: no calculation");
    if (E.calc_pending) {
        HIPCHK(hipEventSynchronize(E.ev_done));
    }
    const int ni = E.last_ni;
    int nslots = 0;
    for (int k = 0; k < ni; ++k)
        if (E.last_h2[k] > 0) E.h_slots[nslots++] = k;
    if (ensure_nb(ni)) return -1;
    E.nb_offset.assign(ni, 0);
    E.nb_valid = true;
    if (nslots == 0) {
        memset(E.h_nbcount, 0, sizeof(int) * (size_t)ni);
        return 0;
    }
    const int ld = (ni + 63) & ~63;
    IBlock ib = iblock_at(E.d_ibuf, ld);
    HIPCHK(hipMemsetAsync(E.d_nbcount, 0, sizeof(int) * (size_t)ni, E.stream));
    HIPCHK(hipMemcpyAsync(E.d_slots, E.h_slots, sizeof(int) * (size_t)nslots,
                          hipMemcpyHostToDevice, E.stream));
    {
        KTimer t(K_NBLIST);
        HIPCHK(launch_nblist(E.stream, E.d_xp64, E.dcap, E.d_jindex, E.last_nj, ib,
                             h2_at(E.d_ibuf, ld), E.d_slots, nslots, E.last_include_self,
                             E.opt.nb_cap, E.d_nbcount, E.d_nblist));
        t.stop();
    }
    HIPCHK(hipMemcpyAsync(E.h_nbcount, E.d_nbcount, sizeof(int) * (size_t)ni,
                          hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    harvest_timing();
    // fetch the stored part of every non-empty list into one host array
    size_t total = 0;
    int overflow = 0;
    for (int s = 0; s < nslots; ++s) {
        const int k = E.h_slots[s];
        const int c = E.h_nbcount[k];
        if (c > E.opt.nb_cap) overflow = 1;
        E.nb_offset[k] = total;
        total += (size_t)std::min(c, E.opt.nb_cap);
    }
    if (total > E.h_nblist_cap) {
        hfree(E.h_nblist);
        E.h_nblist_cap = 0;
        HIPCHK(hipHostMalloc((void **)&E.h_nblist, sizeof(int) * std::max<size_t>(total, 1024),
                             hipHostMallocDefault));
        E.h_nblist_cap = std::max<size_t>(total, 1024);
    }
    for (int s = 0; s < nslots; ++s) {
        const int k = E.h_slots[s];
        const int c = std::min(E.h_nbcount[k], E.opt.nb_cap);
        if (c > 0)
            HIPCHK(hipMemcpyAsync(E.h_nblist + E.nb_offset[k],
                                  E.d_nblist + (size_t)k * E.opt.nb_cap, sizeof(int) * (size_t)c,
                                  hipMemcpyDeviceToHost, E.stream));
    }
    HIPCHK(hipStreamSynchronize(E.stream));
    return overflow;
}

int get_nl(int pipe, int max_length, int *nblen, int *nbl)
{
    if (!E.nb_valid) return fail(-1, "g6_get_neighbour_list: lists not read");
    if (pipe < 0 || pipe >= E.last_ni) {
        *nblen = 0;
        return fail(-1, "g6_get_neighbour_list: pipe %d outside [0, %d)", pipe, E.last_ni);
    }
    const int c = E.h_nbcount[pipe];
    const int stored = std::min(c, E.opt.nb_cap);
    const int n = std::min(stored, std::max(max_length, 0));
    if (n > 0) memcpy(nbl, E.h_nblist + E.nb_offset[pipe], sizeof(int) * (size_t)n);
    if (c >= max_length) {
        *nblen = max_length;
        return 1;
    }
    *nblen = c;
    return c > E.opt.nb_cap ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Autotuner: for each i-block size bucket, time force + reduction (HIP events
// on the library stream) for every candidate shape against the j-memory as
// loaded, and keep the fastest.  The i-block is the first ni j-particles.
// ---------------------------------------------------------------------------
int self_block(int nj, int i_first, int ni, IBlock &ib)
{
    if (ensure_i(ni)) return -1;
    const int ld = (ni + 63) & ~63;
    HIPCHK(launch_iprep_self(E.stream, E.d_jpred, E.d_xp64, E.dcap, i_first, ni,
                             reinterpret_cast<float *>(E.d_ibuf), ld));
    ib = iblock_at(E.d_ibuf, ld);
    ib.x64 = E.d_xp64 + i_first;
    ib.y64 = E.d_xp64 + (size_t)E.dcap + i_first;
    ib.z64 = E.d_xp64 + 2 * (size_t)E.dcap + i_first;
    return 0;
}

// Mean ms of force + reduction for one shape (after one untimed launch).
double time_shape(const Shape &sh, int nj, int ni, const IBlock &ib, float eps2, int reps)
{
    E.forcing_shape = true;
    E.forced = sh;
    Results r = results_at(E.d_res, ni);
    hipEvent_t a = nullptr, b = nullptr;
    double ms_per = -1;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess &&
        enqueue_force(ni, nj, eps2, ib, r) == 0 &&
        hipEventRecord(a, E.stream) == hipSuccess) {
        bool ok = true;
        for (int k = 0; k < reps && ok; ++k) ok = enqueue_force(ni, nj, eps2, ib, r) == 0;
        float ms = 0;
        if (ok && hipEventRecord(b, E.stream) == hipSuccess &&
            hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess)
            ms_per = ms / reps;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    E.forcing_shape = false;
    return ms_per;
}

int autotune(int nj, int ni_max, double eps2)
{
    if (ensure_open("g6x_autotune")) return -1;
    if (E.calc_pending) return fail(-1, "g6x_autotune: a calculation is in flight");
    if (nj <= 0 || nj > E.n_addr) return fail(-1, "g6x_autotune: nj = %d outside j-memory", nj);
    ni_max = std::min(ni_max, nj);
    if (flush_j(nj)) return -1;
    if (predict(nj, true)) return -1;
    const int saved_timing = E.opt.timing;
    E.opt.timing = 0;
    int tuned = 0;
    for (int b = 0; b < kBuckets && (1 << b) <= 2 * ni_max - 1; ++b) {
        const int ni = std::min(1 << b, ni_max);
        IBlock ib;
        if (self_block(nj, 0, ni, ib)) { E.opt.timing = saved_timing; return -1; }
        // about 2 ms of work per candidate, at least 3 launches
        const double est_ms = (double)ni * nj / 1.2e9;
        const int reps = std::max(3, std::min(50, (int)(2.0 / std::max(est_ms, 1e-3))));
        Shape best;
        double best_ms = 1e30;
        const int waves[] = {2048, 4096, 8192, 16384, 32768};
        for (int variant = 0; variant < 2; ++variant) {
            if (variant == 1 && ni < 512) continue;
            for (int tile : {64, 128, kTile}) {
                if (variant == 1 && tile != kTile) continue;
                for (int w : waves) {
                    Shape sh;
                    sh.variant = variant; sh.tile = tile; sh.waves = w; sh.lanes = 0;
                    const double ms = time_shape(sh, nj, ni, ib, (float)eps2, reps);
                    if (ms < 0) { E.opt.timing = saved_timing; return fail(-1, "g6x_autotune: launch failed"); }
                    if (ms < best_ms) { best_ms = ms; best = sh; }
                }
            }
        }
        // reduction lanes for the winning force shape
        Shape lb = best;
        for (int lanes : {4, 8, 16, 32, 64}) {
            Shape sh = best;
            sh.lanes = lanes;
            const double ms = time_shape(sh, nj, ni, ib, (float)eps2, reps);
            if (ms >= 0 && ms < best_ms) { best_ms = ms; lb = sh; }
        }
        E.tuned[b] = lb;
        ++tuned;
    }
    E.opt.timing = saved_timing;
    E.pred_valid = false;
    return tuned;
}

int save_tuning(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return fail(-1, "g6x_save_tuning: cannot write %s", path);
    fprintf(f, "# libg6 force shapes: bucket (ni <= 2^bucket) variant tile waves lanes\n");
    for (int b = 0; b < kBuckets; ++b) {
        const Shape &t = E.tuned[b];
        if (t.waves > 0) fprintf(f, "%d %d %d %d %d\n", b, t.variant, t.tile, t.waves, t.lanes);
    }
    fclose(f);
    return 0;
}

int load_tuning(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(-1, "g6x_load_tuning: cannot read %s", path);
    char line[256];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int b, v, t, w, l;
        if (sscanf(line, "%d %d %d %d %d", &b, &v, &t, &w, &l) != 5) continue;
        if (b < 0 || b >= kBuckets || (v != 0 && v != 1) || w <= 0 ||
            (t != 64 && t != 128 && t != kTile) || (v == 1 && t != kTile) ||
            (l != 0 && l != 4 && l != 8 && l != 16 && l != 32 && l != 64))
            continue;
        Shape &s = E.tuned[b];
        s.variant = v; s.tile = t; s.waves = w; s.lanes = l;
        ++n;
    }
    fclose(f);
    return n;
}

}  // namespace
}  // namespace g6

using namespace g6;

// ===========================================================================
// GRAPE-6 ABI
// ===========================================================================
extern "C" {

void g6_set_tunit_(int *) {}
void g6_set_xunit_(int *) {}

int g6_open_(int *)
{
    read_env();
    return do_open(E.opt.device);
}

int g6_close_(int *) { return do_close(); }

int g6_open_special(int ngpu, int *list)
{
    read_env();
    return do_open(ngpu > 0 && list ? list[0] : E.opt.device);
}

int g6_open_special2(int ngpu, int *list, int)
{
    return g6_open_special(ngpu, list);
}

int g6_reset_(int *)
{
    if (!E.open) return 0;
    (void)hipStreamSynchronize(E.stream);
    E.calc_pending = false;
    E.nb_valid = false;
    E.pred_valid = false;
    return 0;
}

int g6_reset_fofpga_(int *c) { return g6_reset_(c); }

int g6_npipes_(void)
{
    read_env();
    return E.opt.npipes;
}

void g6_set_ti_(int *, double *ti)
{
    E.ti = *ti;
}

int g6_set_j_particle_(int *, int *address, int *index, double *tj, double *dtj,
                       double *mass, double k18[3], double j6[3], double a2[3],
                       double v[3], double x[3])
{
    return set_j(*address, *index, *tj, *dtj, *mass, k18, j6, a2, v, x);
}

void g6calc_firsthalf_(int *, int *nj, int *ni, int index[], double xi[][3],
                       double vi[][3], double[][3], double[][3], double[],
                       double *eps2, double h2[])
{
    (void)firsthalf(*nj, *ni, index, xi, vi, *eps2, h2);
}

int g6calc_lasthalf_(int *, int *, int *ni, int[], double[][3], double[][3],
                     double *, double[], double acc[][3], double jerk[][3],
                     double pot[])
{
    return lasthalf(*ni, acc, jerk, pot, nullptr);
}

int g6calc_lasthalf2_(int *, int *, int *ni, int[], double[][3], double[][3],
                      double *, double[], double acc[][3], double jerk[][3],
                      double pot[], int nnbindex[])
{
    return lasthalf(*ni, acc, jerk, pot, nnbindex);
}

int g6_read_neighbour_list_(int *) { return read_nl(); }
int g6_read_neighbour_list_old_(int *) { return read_nl(); }

int g6_get_neighbour_list_(int *, int *pipe, int *max_length, int *nblen, int nbl[])
{
    return get_nl(*pipe, *max_length, nblen, nbl);
}

int g6_initialize_jp_buffer_(int *, int *) { return 0; }
int g6_flush_jp_buffer_(int *) { return 0; }

// ===========================================================================
// Extensions
// ===========================================================================
int g6x_set_j_particles(int, int n, const int *address, const int *index,
                        const double *tj, const double *dtj, const double *mass,
                        const double *k18, const double *j6, const double *a2,
                        const double *v, const double *x)
{
    for (int k = 0; k < n; ++k) {
        int rc = set_j(address[k], index[k], tj[k], dtj ? dtj[k] : 1.0, mass[k],
                       k18 + 3 * (size_t)k, j6 + 3 * (size_t)k, a2 + 3 * (size_t)k,
                       v + 3 * (size_t)k, x + 3 * (size_t)k);
        if (rc) return rc;
    }
    return 0;
}

int g6x_self_sweep(int, int nj, int i_first, int ni, double ti, double eps2)
{
    if (ensure_open("g6x_self_sweep")) return -1;
    if (ni <= 0 || i_first < 0 || i_first + ni > nj || nj <= 0)
        return fail(-1, "g6x_self_sweep: bad range i_first=%d ni=%d nj=%d", i_first, ni, nj);
    E.calc_pending = false;
    E.nb_valid = false;
    E.ti = ti;
    if (flush_j(nj)) return -1;
    if (predict(nj, true)) return -1;
    IBlock ib;
    if (self_block(nj, i_first, ni, ib)) return -1;
    Results r = results_at(E.d_res, ni);
    if (enqueue_force(ni, nj, (float)eps2, ib, r)) return -1;
    E.last_ni = ni;
    E.last_nj = nj;
    E.calc_failed = true;   // no neighbour lists for sweeps
    return 0;
}

int g6x_read_results(int, int ni, double *acc, double *jerk, double *pot, int *nn)
{
    if (ensure_open("g6x_read_results")) return -1;
    if (ni > E.last_ni) return fail(-1, "g6x_read_results: ni %d > last %d", ni, E.last_ni);
    HIPCHK(hipMemcpyAsync(E.h_res, E.d_res, results_bytes(E.last_ni), hipMemcpyDeviceToHost,
                          E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    harvest_timing();
    Results h = results_at(E.h_res, E.last_ni);
    E.last_refinements = count_refined(h, E.last_ni);
    if (acc) memcpy(acc, h.acc, sizeof(double) * 3 * (size_t)ni);
    if (jerk) memcpy(jerk, h.jerk, sizeof(double) * 3 * (size_t)ni);
    if (pot) memcpy(pot, h.pot, sizeof(double) * (size_t)ni);
    if (nn) memcpy(nn, h.nn, sizeof(int) * (size_t)ni);
    return 0;
}

int g6x_read_predicted(int, int nj, double *xp, double *vp)
{
    if (ensure_open("g6x_read_predicted")) return -1;
    if (!E.pred_valid || nj > E.pred_nj) return fail(-1, "g6x_read_predicted: no prediction");
    std::vector<double> buf(3 * (size_t)nj);
    for (int pass = 0; pass < 2; ++pass) {
        const double *src = pass ? E.d_vp64 : E.d_xp64;
        double *dst = pass ? vp : xp;
        if (!dst) continue;
        for (int c = 0; c < 3; ++c) {
            HIPCHK(hipMemcpyAsync(buf.data() + (size_t)c * nj, src + (size_t)c * E.dcap,
                                  sizeof(double) * (size_t)nj, hipMemcpyDeviceToHost, E.stream));
        }
        HIPCHK(hipStreamSynchronize(E.stream));
        for (int j = 0; j < nj; ++j)
            for (int c = 0; c < 3; ++c) dst[3 * (size_t)j + c] = buf[(size_t)c * nj + j];
    }
    return 0;
}

int g6x_sync(int)
{
    if (!E.open) return 0;
    HIPCHK(hipStreamSynchronize(E.stream));
    HIPCHK(hipGetLastError());
    harvest_timing();
    return 0;
}

int g6x_set_option(const char *key, long long value)
{
    read_env();
    std::string k(key ? key : "");
    Options &o = E.opt;
    if (k == "predict_order") o.predict_order = (int)value;
    else if (k == "npipes") o.npipes = std::max(1, (int)value);
    else if (k == "nb_cap") o.nb_cap = std::max(1, (int)value);
    else if (k == "waves_target") o.waves_target = std::max(1, (int)value);
    else if (k == "nsplit") o.nsplit = std::max(0, (int)value);
    else if (k == "timing") o.timing = (int)value;
    else if (k == "xcd_remap") o.xcd_remap = (int)value;
    else if (k == "force_variant") o.force_variant = (int)value;
    else if (k == "lds_min_ni") o.lds_min_ni = (int)value;
    else if (k == "tile") o.tile = (int)value;
    else if (k == "reduce_lanes") {
        if (value != 0 && value != 4 && value != 8 && value != 16 && value != 32 && value != 64)
            return -1;
        o.reduce_lanes = (int)value;
    }
    else return -1;
    E.pred_valid = false;
    return 0;
}

long long g6x_get_option(const char *key)
{
    read_env();
    std::string k(key ? key : "");
    const Options &o = E.opt;
    if (k == "predict_order") return o.predict_order;
    if (k == "npipes") return o.npipes;
    if (k == "nb_cap") return o.nb_cap;
    if (k == "waves_target") return o.waves_target;
    if (k == "nsplit") return o.nsplit;
    if (k == "timing") return o.timing;
    if (k == "xcd_remap") return o.xcd_remap;
    if (k == "force_variant") return o.force_variant;
    if (k == "lds_min_ni") return o.lds_min_ni;
    if (k == "tile") return o.tile;
    if (k == "last_tile") return E.geom_tile;
    if (k == "reduce_lanes") return o.reduce_lanes;
    return -1;
}

int g6x_kernel_time(int kernel, double *ms)
{
    if (kernel < 0 || kernel >= K_COUNT) return -1;
    if (ms) *ms = E.k_ms[kernel];
    return E.k_calls[kernel];
}

void g6x_reset_timing(void)
{
    for (int k = 0; k < K_COUNT; ++k) { E.k_ms[k] = 0; E.k_calls[k] = 0; }
}

int g6x_last_force_geometry(int *workgroups, int *nsplit, int *jchunk)
{
    if (workgroups) *workgroups = E.geom_wg;
    if (nsplit) *nsplit = E.geom_nsplit;
    if (jchunk) *jchunk = E.geom_jchunk;
    return E.geom_variant;
}

int g6x_last_refinements(void) { return E.last_refinements; }

int g6x_autotune(int, int nj, int ni_max, double eps2)
{
    read_env();
    return autotune(nj, ni_max, eps2);
}

int g6x_get_tuning(int bucket, int *variant, int *tile, int *waves, int *lanes)
{
    if (bucket < 0 || bucket >= kBuckets) return -1;
    const Shape &t = E.tuned[bucket];
    if (variant) *variant = t.variant;
    if (tile) *tile = t.tile;
    if (waves) *waves = t.waves;
    if (lanes) *lanes = t.lanes;
    return t.waves > 0 ? 1 : 0;
}

int g6x_clear_tuning(void)
{
    for (Shape &t : E.tuned) t = Shape();
    return 0;
}

int g6x_save_tuning(const char *path) { return save_tuning(path ? path : ""); }

int g6x_load_tuning(const char *path) { return load_tuning(path ? path : ""); }

const char *g6x_last_error(void) { return E.last_error.c_str(); }

}  // extern "C"
