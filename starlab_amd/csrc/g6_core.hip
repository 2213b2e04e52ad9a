// g6_core.hip -- the engine of libg6 / libg4x on MI355X (gfx950): j-memory
// (pinned host mirror, HBM copy, dirty tracking), the j predictor trigger,
// i-block staging, force geometry and launch, neighbour lists, timing and the
// force-shape autotuner.  The GRAPE-6 ABI (g6_engine.hip) and the GRAPE-4 ABI
// (g4_engine.hip) are thin front ends over these functions.
//
// The call sequence kira uses with GRAPE-6 (hdyn_grape6.C:474-910,
// 1991-2181, 3067-3411) is open -> set_j* -> set_ti -> firsthalf ->
// lasthalf2 -> [read_nl -> get_nl*]; with GRAPE-4 (hdyn_grape4.C:240-420)
// it is open -> jpdma* -> setti -> calc_firsthalf -> calc_lasthalf ->
// nbread -> nblist*.  Both repeat, with close / re-open and resets at any
// point.
#include "g6_core.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace g6 {

Engine E;

int bucket_of(int ni)
{
    int b = 0;
    while (b < kBuckets - 1 && (1 << b) < ni) ++b;
    return b;
}

// ---------------------------------------------------------------------------
int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    E.last_error = buf;
    fprintf(stderr, "%s: %s\n", lib_tag(), buf);
    return code;
}

long long env_ll(const char *name, long long dflt)
{
    const std::string key = std::string(env_prefix()) + name;
    const char *s = getenv(key.c_str());
    if (!s || !*s) return dflt;
    return atoll(s);
}

// Environment knobs, named <prefix><knob> with the front end's prefix
// (G6_NPIPES for libg6, G4_NPIPES for libg4x, ...).
void read_env()
{
    if (E.env_read) return;
    E.env_read = true;
    Options &o = E.opt;
    o.predict_order = (int)env_ll("PREDICT_ORDER", o.predict_order);
    o.npipes = (int)env_ll("NPIPES", o.npipes);
    o.nb_cap = (int)env_ll("NB_CAP", o.nb_cap);
    o.waves_target = (int)env_ll("WAVES", o.waves_target);
    o.nsplit = (int)env_ll("NSPLIT", o.nsplit);
    o.timing = (int)env_ll("TIMING", o.timing);
    o.device = (int)env_ll("DEVICE", o.device);
    o.xcd_remap = (int)env_ll("XCD_REMAP", o.xcd_remap);
    o.force_variant = (int)env_ll("FORCE_VARIANT", o.force_variant);
    o.lds_min_ni = (int)env_ll("LDS_MIN_NI", o.lds_min_ni);
    o.tile = (int)env_ll("TILE", o.tile);
    o.reduce_lanes = (int)env_ll("REDUCE_LANES", o.reduce_lanes);
    o.wpg = (int)env_ll("WPG", o.wpg);
    o.fused_small = (int)env_ll("FUSED_SMALL", o.fused_small);
    o.small_fp64 = (int)env_ll("SMALL_FP64", o.small_fp64);
    o.done_word = (int)env_ll("DONE_WORD", o.done_word);
    o.fail_lasthalf = (int)env_ll("FAIL_LASTHALF", o.fail_lasthalf);
    o.fail_read_nl = (int)env_ll("FAIL_READNL", o.fail_read_nl);
    o.fail_nn = (int)env_ll("FAIL_NN", o.fail_nn);
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

int wait_queued(unsigned long long seq);
unsigned long long mark_queued();

int grow_host_mirror(int need)
{
    if (need <= E.hcap) return 0;
    // a queued scatter or upload may still read the old mirror
    if (wait_queued(E.mirror_seq)) return -1;
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
    E.h_jraw_dev = nullptr;
    E.hcap = cap;
    E.dirty_flag.resize(cap, 0);
    void *dev = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dev, nr, 0));
    E.h_jraw_dev = static_cast<const JRaw *>(dev);
    return 0;
}

void free_device_j()
{
    dfree(E.d_jraw); dfree(E.d_jpred); dfree(E.d_xp64); dfree(E.d_vp64); dfree(E.d_jindex);
    E.dcap = 0;
    E.pred_valid = false;
    E.xp_valid = false;
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
    // vp64 [3][cap] followed by m64 [cap]
    HIPCHK(hipMalloc((void **)&E.d_vp64, sizeof(double) * 4 * (size_t)cap));
    HIPCHK(hipMalloc((void **)&E.d_jindex, sizeof(int) * (size_t)cap));
    HIPCHK(hipMemsetAsync(E.d_jraw, 0, sizeof(JRaw) * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_jpred, 0, sizeof(JPred) * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_xp64, 0, sizeof(double) * 3 * (size_t)cap, E.stream));
    HIPCHK(hipMemsetAsync(E.d_vp64, 0, sizeof(double) * 4 * (size_t)cap, E.stream));
    E.dcap = cap;
    // everything set so far must reach the new arrays (a deferred scatter's
    // records are in the mirror too)
    for (int a : E.dirty) E.dirty_flag[a] = 0;
    E.dirty.clear();
    E.deferred_scatter = 0;
    const int n = std::min(E.n_addr, E.hcap);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(E.d_jraw, E.h_jraw, sizeof(JRaw) * (size_t)n,
                              hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(E.d_jindex, E.h_jindex, sizeof(int) * (size_t)n,
                              hipMemcpyHostToDevice, E.stream));
        E.mirror_seq = mark_queued();
    }
    return 0;
}

// Wait until the queued op numbered seq (one that reads or writes a pinned
// buffer) has completed.  In kira's call order lasthalf has already waited
// for it, so this costs nothing.
int wait_queued(unsigned long long seq)
{
    if (seq <= E.synced_seq) return 0;
    HIPCHK(hipStreamSynchronize(E.stream));
    E.synced_seq = E.op_seq;
    return 0;
}

unsigned long long mark_queued() { return ++E.op_seq; }

int ensure_stage(int n)
{
    if (n <= E.scap) return 0;
    int cap = std::max(n, std::max(256, E.scap * 2));
    if (wait_queued(E.stage_seq)) return -1;
    hfree(E.h_stage);
    E.h_stage_dev = nullptr;
    E.scap = 0;
    const size_t bytes = 2 * sizeof(int) * (size_t)cap;   // addresses, then index fields
    HIPCHK(hipHostMalloc((void **)&E.h_stage, bytes, hipHostMallocDefault));
    void *dev = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dev, E.h_stage, 0));
    E.h_stage_dev = static_cast<const char *>(dev);
    E.scap = cap;
    return 0;
}

// Push dirty j records to HBM: the whole used range when many changed,
// otherwise a packed list scattered by a kernel.
// Launch a scatter that calc_begin packed but did not get to launch.
int launch_deferred_scatter()
{
    const int nd = E.deferred_scatter;
    if (nd <= 0 || E.dcap == 0) return 0;
    E.deferred_scatter = 0;
    KTimer t(K_SCATTER);
    HIPCHK(launch_scatter(E.stream, E.d_jraw, E.d_jindex, E.h_jraw_dev,
                          reinterpret_cast<const int *>(E.h_stage_dev), nd));
    t.stop();
    E.stage_seq = E.mirror_seq = mark_queued();
    E.pred_valid = false;
    return 0;
}

int flush_j(int nj, bool defer, bool noalloc)
{
    if (launch_deferred_scatter()) return -1;
    if (noalloc) {
        // only copies and kernels: no buffer may grow
        if (nj > E.n_addr)
            return fail(-1, "j-memory holds %d addresses, not %d (set them first)", E.n_addr, nj);
        if (E.n_addr > E.dcap)
            return fail(-1, "j-memory is not resident on the device yet (any force call, "
                            "sweep or g6x_autotune after the last new address places it)");
    }
    if (nj > E.n_addr) {
        // addresses never set are zero-mass particles at the origin
        if (grow_host_mirror(nj)) return -1;
        E.n_addr = nj;
    }
    if (ensure_device_j(E.n_addr)) return -1;
    const int nd = (int)E.dirty.size();
    if (nd == 0) return 0;
    // without allocation a dirty list longer than the staging buffer goes
    // up as one copy of the whole used range
    if ((size_t)nd * 4 >= (size_t)E.n_addr || (noalloc && nd > E.scap)) {
        HIPCHK(hipMemcpyAsync(E.d_jraw, E.h_jraw, sizeof(JRaw) * (size_t)E.n_addr,
                              hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(E.d_jindex, E.h_jindex, sizeof(int) * (size_t)E.n_addr,
                              hipMemcpyHostToDevice, E.stream));
        E.mirror_seq = mark_queued();
    } else {
        // the staging buffer may still be read by an earlier scatter
        if (wait_queued(E.stage_seq) || ensure_stage(nd)) return -1;
        // the list k_scatter reads from pinned memory: addresses, then their
        // index fields; the records themselves it reads from the mirror
        int *list = reinterpret_cast<int *>(E.h_stage);
        for (int k = 0; k < nd; ++k) {
            const int a = E.dirty[k];
            list[k] = a;
            list[nd + k] = E.h_jindex[a];
        }
        if (defer) {
            E.deferred_scatter = nd;
        } else {
            KTimer t(K_SCATTER);
            HIPCHK(launch_scatter(E.stream, E.d_jraw, E.d_jindex, E.h_jraw_dev,
                                  reinterpret_cast<const int *>(E.h_stage_dev), nd));
            t.stop();
            E.stage_seq = E.mirror_seq = mark_queued();
        }
    }
    for (int a : E.dirty) E.dirty_flag[a] = 0;
    E.dirty.clear();
    // the device j-memory changed: neither prediction describes it any more
    E.pred_valid = false;
    E.xp_valid = false;
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
    E.xp_valid = true;
    E.xp_ti = E.ti;
    E.xp_nj = nj;
    return 0;
}

// i-block buffer: 10 four-byte arrays (xh yh zh xl yl zl vx vy vz excl), then
// 8 double arrays (x64 y64 z64 h2 e2 vx64 vy64 vz64), each of stride ld.
size_t ibuf_bytes(int cap) { return (size_t)cap * (10 * 4 + 8 * 8); }

// results: 7 doubles, nn and the refinement flag per i
size_t results_bytes(int ni) { return (size_t)ni * (7 * sizeof(double) + 2 * sizeof(int)); }

int ensure_i(int ni)
{
    if (ni <= E.icap) return 0;
    int cap = std::max((ni + 63) & ~63, 256);
    if (wait_queued(std::max(E.ibuf_seq, E.res_seq))) return -1;
    hfree(E.h_ibuf); dfree(E.d_ibuf); hfree(E.h_slots);
    hfree(E.h_res); dfree(E.d_res);
    E.h_res_dev = nullptr;
    E.h_ibuf_dev = nullptr;
    E.icap = 0;
    HIPCHK(hipHostMalloc((void **)&E.h_ibuf, ibuf_bytes(cap), hipHostMallocDefault));
    {
        void *dev = nullptr;
        HIPCHK(hipHostGetDevicePointer(&dev, E.h_ibuf, 0));
        E.h_ibuf_dev = static_cast<const char *>(dev);
    }
    HIPCHK(hipMalloc((void **)&E.d_ibuf, ibuf_bytes(cap)));
    HIPCHK(hipHostMalloc((void **)&E.h_slots, sizeof(int) * (size_t)cap, hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer((void **)&E.h_slots_dev, E.h_slots, 0));
    const size_t rb = results_bytes(cap);
    HIPCHK(hipHostMalloc((void **)&E.h_res, rb, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&E.d_res, rb));
    void *dev = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dev, E.h_res, 0));
    E.h_res_dev = static_cast<char *>(dev);
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
    b.vx64 = d + 5 * (size_t)ld; b.vy64 = d + 6 * (size_t)ld; b.vz64 = d + 7 * (size_t)ld;
    b.ld = ld;
    return b;
}

double *h2_at(char *base, int ld)
{
    return reinterpret_cast<double *>(base + (size_t)ld * 40) + 3 * (size_t)ld;
}

double *e2_at(char *base, int ld)
{
    return reinterpret_cast<double *>(base + (size_t)ld * 40) + 4 * (size_t)ld;
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
    // grow geometrically, so blocks of slowly rising size reallocate rarely
    const size_t b = std::max(bytes, E.pbytes * 2);
    dfree(E.d_part);
    E.pbytes = 0;
    HIPCHK(hipMalloc((void **)&E.d_part, b));
    E.pbytes = b;
    return 0;
}

Partials partials_at(int ld, int nsplit, char *base)
{
    Partials p;
    const size_t n = (size_t)ld * nsplit;
    double *d = reinterpret_cast<double *>(base ? base : E.d_part);
    for (int c = 0; c < 7; ++c) p.c[c] = d + c * n;
    float *u = reinterpret_cast<float *>(d + 7 * n);
    p.best = u;
    p.btile = reinterpret_cast<int *>(u + n);
    p.second = u + 2 * n;
    p.tsum = u + 3 * n;
    p.tstride = n;
    p.ld = ld;
    return p;
}

// per (split, i): 7 FP64 sums, best, tile, runner-up, 10 FP32 tile sums
size_t partials_bytes(int ld, int nsplit) { return (size_t)ld * nsplit * (7 * 8 + 13 * 4); }

struct Geometry {
    int variant, tile, nsplit, jchunk, lanes;
    int wpg;        // waves per workgroup (scalar-load kernel)
    int npart;      // partial splits the reduction reads
};

bool valid_wpg(int w) { return w == 1 || w == 2 || w == 4 || w == 8 || w == 16; }

// waves of the small-block force shape: one 16-wave workgroup per CU; used
// up to 4 i-groups (ni <= 256), beyond which it measured slower
constexpr int kSmallWaves = 4096;
constexpr int kSmallMaxGroups = 4;

// j per FP32 tile of the scalar-load kernel: 16 and 32 shorten each wave's
// chain of dependent scalar loads for tiny i-blocks, 256 amortises the FP64
// folds of big ones
bool valid_tile(int t) { return t == 16 || t == 32 || t == 64 || t == 128 || t == kTile; }

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
        if (valid_tile(E.opt.tile)) sh.tile = E.opt.tile;
        if (E.opt.reduce_lanes > 0) sh.lanes = E.opt.reduce_lanes;
        if (valid_wpg(E.opt.wpg)) sh.wpg = E.opt.wpg;
        if (t.waves == 0 || E.opt.waves_target != Options().waves_target)
            sh.waves = E.opt.waves_target;
    }
    Geometry g;
    g.variant = sh.variant;
    const int per_wave = force_igroup(g.variant) / (g.variant == 1 ? 4 : 1);
    const int n_iw = std::max(1, (ni + per_wave - 1) / per_wave);
    const int target = std::max(1, sh.waves);
    const bool nsplit_set = !E.forcing_shape && E.opt.nsplit > 0;
    // Small i-blocks (at most 4 i-groups): one 16-wave workgroup per CU,
    // kSmallWaves waves in all, each on a short chain of j (16..256 per tile).
    // Measured through the ABI at ni <= 256: 2-14% less time per block than
    // 64-j tiles with 8-wave workgroups.
    bool small = false;
    if (g.variant == 1) {
        g.tile = kTile;
    } else if (valid_tile(sh.tile)) {
        g.tile = sh.tile;
    } else if ((long long)n_iw * ((nj + kTile - 1) / kTile) >= target) {
        g.tile = kTile;
    } else if (!nsplit_set && n_iw <= kSmallMaxGroups) {
        small = true;
        const int want = std::max(1, kSmallWaves / n_iw);
        g.tile = 16;
        while (g.tile < kTile && (long long)g.tile * want < nj) g.tile *= 2;
    } else {
        g.tile = 64;
    }
    const int ntiles = std::max(1, (nj + g.tile - 1) / g.tile);
    int sp = nsplit_set ? E.opt.nsplit
                        : std::max(1, ((small ? kSmallWaves : target) + n_iw - 1) / n_iw);
    sp = std::min(sp, ntiles);
    const int tiles_per = (ntiles + sp - 1) / sp;
    g.jchunk = tiles_per * g.tile;
    g.nsplit = std::max(1, (nj + g.jchunk - 1) / g.jchunk);
    // waves per workgroup: the splits of a workgroup fold their sums in LDS;
    // built-in rule 16 for the small-block shape, else 8 for many splits and
    // 4 otherwise, at most nsplit (measured at N = 65536: 16 waves per
    // workgroup is slower for large blocks)
    if (g.variant == 1) {
        g.wpg = 1;
    } else if (valid_wpg(sh.wpg)) {
        g.wpg = sh.wpg;
    } else if (small) {
        g.wpg = 1;
        while (2 * g.wpg <= std::min(16, g.nsplit)) g.wpg *= 2;
    } else {
        const int cap = g.nsplit >= 64 ? 8 : 4;
        g.wpg = 1;
        while (2 * g.wpg <= std::min(cap, g.nsplit)) g.wpg *= 2;
    }
    g.npart = force_partials(g.variant, g.nsplit, g.wpg);
    g.lanes = sh.lanes > 0 ? sh.lanes : reduce_lanes(g.npart);
    return g;
}

// Enqueue force + reduction (with the nearest-neighbour refinement) for an
// i-block already in d_ibuf.
int enqueue_force(int ni, int nj, double eps2d, const IBlock &ib, const Results &r, bool zskip,
                  const FusedPredict *fused, char *part, const DoneWord *done)
{
    float eps2 = (float)eps2d;
    if (zskip && !(eps2 >= __FLT_MIN__)) {   // as launch_force treats it
        eps2 = 0.0f;
        eps2d = 0.0;
    }
    const Geometry g = choose_geometry(ni, nj);
    const int ld = ib.ld;
    if (!part && ensure_partials(partials_bytes(ld, g.npart))) return -1;
    Partials p = partials_at(ld, g.npart, part);
    E.geom_nsplit = g.nsplit;
    E.geom_jchunk = g.jchunk;
    E.geom_tile = g.tile;
    E.geom_lanes = g.lanes;
    E.geom_wpg = g.wpg;
    E.geom_wg = ((ni + force_igroup(g.variant) - 1) / force_igroup(g.variant)) * g.npart;
    E.geom_variant = g.variant;
    {
        KTimer t(K_FORCE);
        HIPCHK(launch_force(E.stream, E.d_jpred, nj, ib, ni, eps2, g.nsplit, g.jchunk, g.tile,
                            p, E.opt.xcd_remap != 0, g.variant, zskip, g.wpg, fused));
        t.stop();
    }
    {
        KTimer t(K_REDUCE);
        const NearTile nt{eps2d, g.jchunk, g.variant == 1 ? 1 : g.wpg, zskip ? 1 : 0,
                          E.d_vp64, E.d_vp64 + 3 * (size_t)E.dcap};
        HIPCHK(launch_reduce(E.stream, p, g.npart, ni, nj, ib, eps2, g.tile,
                             E.d_xp64, E.dcap, E.d_jindex, r, g.lanes, nt, done));
        t.stop();
    }
    return 0;
}

int ensure_open(const char *who)
{
    if (!E.open) return fail(-1, "%s: device not open", who);
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
        return fail(-1, "open: no HIP device available (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "0 devices");
    if (dev < 0 || dev >= n) dev = 0;
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreate(&E.stream));
    const std::string tune_env = std::string(env_prefix()) + "TUNE_FILE";
    if (const char *tf = getenv(tune_env.c_str())) {
        if (*tf && load_tuning(tf) < 0) E.last_error.clear();   // a missing file is not an error
    }
    HIPCHK(hipEventCreateWithFlags(&E.ev_done, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void **)&E.h_flag, 64, hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void **)&E.h_flag_dev, E.h_flag, 0));
    __atomic_store_n(E.h_flag, 0ull, __ATOMIC_RELEASE);
    HIPCHK(hipMalloc((void **)&E.d_done_ticket, 256));
    HIPCHK(hipMemset(E.d_done_ticket, 0, 256));
    E.flag_seq = 0;
    E.flag_pending = false;
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
    hfree(E.h_stage);
    E.h_stage_dev = nullptr;
    E.scap = 0;
    hfree(E.h_ibuf); dfree(E.d_ibuf); hfree(E.h_slots);
    hfree(E.h_res); dfree(E.d_res);
    E.h_res_dev = nullptr;
    E.h_ibuf_dev = nullptr;
    E.deferred_scatter = 0;   // the next open uploads the whole mirror
    E.synced_seq = E.op_seq;
    E.icap = 0;
    dfree(E.d_part); E.pbytes = 0;
    dfree(E.d_small);
    dfree(E.d_done_ticket);
    dfree(E.d_nblist); hfree(E.h_nbcount); hfree(E.h_nblist);
    dfree(E.d_nlscratch); E.nlscratch_cap = 0;
    hfree(E.h_nlinfo); E.nlinfo_cap = 0;
    E.nbcap_entries = 0; E.h_nblist_cap = 0; E.nbcount_cap = 0;
    for (hipEvent_t ev : E.ev_pool) (void)hipEventDestroy(ev);
    E.ev_pool.clear();
    if (E.ev_done) (void)hipEventDestroy(E.ev_done);
    E.ev_done = nullptr;
    hfree(E.h_flag);
    E.h_flag_dev = nullptr;
    E.flag_pending = false;
    if (E.stream) (void)hipStreamDestroy(E.stream);
    E.stream = nullptr;
    E.open = false;
    // no calculation survives a close: lasthalf and the neighbour-list calls
    // report an error until the next firsthalf
    E.calc_pending = false;
    E.calc_failed = false;
    E.last_ni = 0;
    E.sweep_results = false;
    E.nb_valid = false;
    return 0;
}

void index_map_set(int index, int address);

int set_j(int address, int index, double tj, double dtj, double mass,
          const double *k18, const double *j6, const double *a2,
          const double *v, const double *x)
{
    if (address < 0) return fail(-1, "g6_set_j_particle: negative address %d", address);
    // a queued scatter reads records from the mirror: a block in flight keeps
    // the j state it was started with (no wait in kira's call order, where
    // lasthalf has already synchronised)
    if (wait_queued(E.mirror_seq)) return -1;
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
        if (excl_address(old) == address) index_map_set(old, -1);
        E.h_jindex[address] = index;
        index_map_set(index, address);
    } else if (excl_address(index) != address) {
        index_map_set(index, address);   // the index was taken by another address since
    }
    if (address + 1 > E.n_addr) E.n_addr = address + 1;
    if (!E.dirty_flag[address]) {
        E.dirty_flag[address] = 1;
        E.dirty.push_back(address);
    }
    return 0;
}

// A reset drops the calculation in flight and forces a fresh prediction.
void reset_state()
{
    if (!E.open) return;
    if (hipStreamSynchronize(E.stream) == hipSuccess) E.synced_seq = E.op_seq;
    E.calc_pending = false;
    E.nb_valid = false;
    E.pred_valid = false;
}

// Index fields of kira's nodes are small non-negative integers: they map
// through a flat array; anything else through a hash map.
constexpr int kDenseIndex = 1 << 24;

int excl_address(int index)
{
    if (index >= 0 && index < kDenseIndex)
        return (size_t)index < E.addr_by_index.size() ? E.addr_by_index[index] : -1;
    auto it = E.addr_of_index.find(index);
    return it == E.addr_of_index.end() ? -1 : it->second;
}

// address < 0 removes the entry
void index_map_set(int index, int address)
{
    if (index >= 0 && index < kDenseIndex) {
        if ((size_t)index >= E.addr_by_index.size()) {
            if (address < 0) return;
            E.addr_by_index.resize(std::max<size_t>((size_t)index + 1,
                                                    2 * E.addr_by_index.size()), -1);
        }
        E.addr_by_index[index] = address;
    } else if (address < 0) {
        E.addr_of_index.erase(index);
    } else {
        E.addr_of_index[index] = address;
    }
}

// Stage the i-block in the pinned i buffer (layout of ibuf_bytes(ld)) and
// return its stride ld, or -1.
int stage_iblock(int ni, const double (*xi)[3], const double (*vi)[3], const int *excl,
                 const double *h2, const double *e2)
{
    // the pinned i buffer may still be read by the previous upload
    if (wait_queued(E.ibuf_seq) || ensure_i(ni)) return -1;
    const int ld = (ni + 63) & ~63;
    float *f = reinterpret_cast<float *>(E.h_ibuf);
    int *ex = reinterpret_cast<int *>(f + 9 * (size_t)ld);
    double *d = reinterpret_cast<double *>(E.h_ibuf + (size_t)ld * 40);
    double *hh = d + 3 * (size_t)ld;
    double *ee = d + 4 * (size_t)ld;
    for (int k = 0; k < ld; ++k) {
        const int s = k < ni ? k : ni - 1;
        for (int c = 0; c < 3; ++c) {
            const double x = xi[s][c];
            const float hi = (float)x;
            f[(0 + c) * (size_t)ld + k] = hi;
            f[(3 + c) * (size_t)ld + k] = (float)(x - (double)hi);
            f[(6 + c) * (size_t)ld + k] = (float)vi[s][c];
            d[c * (size_t)ld + k] = x;
            d[(5 + c) * (size_t)ld + k] = vi[s][c];
        }
        ex[k] = (k < ni && excl) ? excl[k] : -1;
        hh[k] = (k < ni && h2) ? h2[k] : 0.0;
        ee[k] = (k < ni && e2) ? e2[k] : 0.0;
    }
    return ld;
}

// One launch: the deferred j scatter and the upload of the staged i-block
// (ld > 0; ld = 0 uploads nothing).
int scatter_with_iblock(int ld)
{
    const int nd = E.deferred_scatter;
    if (nd <= 0 && ld <= 0) return 0;
    E.deferred_scatter = 0;
    KTimer t(K_SCATTER);
    HIPCHK(launch_scatter(E.stream, E.d_jraw, E.d_jindex, E.h_jraw_dev,
                          reinterpret_cast<const int *>(E.h_stage_dev), nd, E.h_ibuf_dev,
                          E.d_ibuf, ld > 0 ? ibuf_bytes(ld) : 0));
    t.stop();
    const unsigned long long seq = mark_queued();
    if (ld > 0) E.ibuf_seq = seq;
    if (nd > 0) E.stage_seq = E.mirror_seq = seq;
    return 0;
}

// A block of at most kSmallUseI i with one softening and no zero-distance
// rule: one k_small launch (FP64 pairs, the reduction inside).  Pending j
// updates travel in its arguments when they fit and lie below nj; otherwise
// the scatter launch runs first, as for other blocks.  With neighbour
// spheres (lists) the i-block is uploaded too and k_small writes the FP64
// prediction, both of which the list kernels read.
int calc_small(int nj, int ni, const double (*xi)[3], const double (*vi)[3], const int *excl,
               const double *h2, const double *e2, double eps2, bool lists)
{
    // the pinned results buffer (and the i buffer) come with the i capacity
    if (ensure_i(ni)) return -1;
    if (!E.h_res_dev) return fail(-1, "firsthalf: no results buffer");
    const size_t npart = small_part_doubles();
    if (!E.d_small) {
        HIPCHK(hipMalloc((void **)&E.d_small, sizeof(double) * npart + 256));
        HIPCHK(hipMemsetAsync(E.d_small + npart, 0, 256, E.stream));
    }
    static SmallBlock B;   // the full form; launch_small passes the smallest form that holds it
    B.ti = E.ti;
    B.eps2 = eps2;
    B.ni = ni;
    B.nj = nj;
    B.order = E.opt.predict_order;
    B.write_pred = lists ? 1 : 0;
    B.nd = 0;
    for (int k = 0; k < ni; ++k) {
        for (int c = 0; c < 3; ++c) {
            B.xi[k][c] = xi[k][c];
            B.vi[k][c] = vi[k][c];
        }
        B.excl[k] = excl ? excl[k] : -1;
    }
    const int nd = E.deferred_scatter;
    const int *list = reinterpret_cast<const int *>(E.h_stage);
    bool inline_j = !lists && nd <= kSmallMaxDirty;
    for (int k = 0; k < nd && inline_j; ++k) inline_j = list[k] < nj;
    if (nd > 0 && inline_j) {
        for (int k = 0; k < nd; ++k) {
            B.daddr[k] = list[k];
            B.dindex[k] = list[nd + k];
            B.drec[k] = E.h_jraw[list[k]];
        }
        B.nd = nd;
        E.deferred_scatter = 0;
    } else {
        const int ld = lists ? stage_iblock(ni, xi, vi, excl, h2, e2) : 0;
        if (ld < 0 || scatter_with_iblock(ld)) return -1;
    }
    const Results r = results_at(E.h_res_dev, ni);
    KTimer t(K_FORCE);
    // with kernel timing on, lasthalf waits for the event (it reads the timers)
    const bool flag = E.h_flag_dev && E.opt.done_word && !E.opt.timing;
    const hipError_t e = launch_small(E.stream, B, E.d_jraw, E.d_jindex, E.d_xp64, E.d_vp64,
                                      E.dcap, E.d_small,
                                      reinterpret_cast<unsigned *>(E.d_small + npart), r,
                                      flag ? E.h_flag_dev : nullptr, E.flag_seq + 1);
    t.stop();
    if (e == hipSuccess && flag) {
        E.flag_seq += 1;
        E.flag_pending = true;
    }
    if (e != hipSuccess) {
        // updates carried in the arguments never reached HBM: send them again
        for (int k = 0; k < B.nd; ++k) {
            const int a = B.daddr[k];
            if (!E.dirty_flag[a]) {
                E.dirty_flag[a] = 1;
                E.dirty.push_back(a);
            }
        }
        return fail(-1, "k_small launch: %s", hipGetErrorString(e));
    }
    if (B.nd > 0) E.pred_valid = false;
    E.xp_valid = lists;
    if (lists) {
        E.xp_ti = E.ti;
        E.xp_nj = nj;
    }
    int wgs, chunks;
    small_grid(nj, wgs, chunks);
    E.geom_variant = 2;
    E.geom_nsplit = wgs;
    E.geom_jchunk = kSmallThreads * chunks;
    E.geom_wg = wgs;
    E.geom_tile = 0;
    E.geom_wpg = 0;
    E.geom_lanes = 64;
    return 0;
}

// Largest FP32-path block whose reduction stores the completion word.
constexpr int kDoneWordMaxI = 512;

int calc_begin(const char *who, int nj, int ni, const double (*xi)[3], const double (*vi)[3],
               const int *excl, const double *h2, const double *e2, const IGroup *groups,
               int ngroups, int include_self)
{
    E.calc_pending = false;
    E.calc_failed = true;
    E.sweep_results = false;
    E.nb_valid = false;
    E.flag_pending = false;
    if (ensure_open(who)) return -1;
    if (ni <= 0) return fail(-1, "%s: ni = %d", who, ni);
    if (nj <= 0) return fail(-1, "%s: nj = %d", who, nj);
    int covered = 0;
    for (int g = 0; g < ngroups; ++g) {
        if (groups[g].off != covered || groups[g].n <= 0)
            return fail(-1, "%s: i-groups do not tile the block", who);
        covered += groups[g].n;
    }
    if (covered != ni) return fail(-1, "%s: i-groups cover %d of %d slots", who, covered, ni);
    // the j scatter is launched below, together with the i-block upload
    if (flush_j(nj, true)) return -1;
    // no h2 array: no neighbour spheres
    if (h2) E.last_h2.assign(h2, h2 + ni);
    else E.last_h2.assign(ni, 0.0);
    bool lists = false;
    for (int k = 0; k < ni && h2; ++k) lists = lists || h2[k] > 0;
    if (E.opt.small_fp64 && ngroups == 1 && !groups[0].zskip && ni <= kSmallUseI &&
        nj <= E.dcap) {
        if (calc_small(nj, ni, xi, vi, excl, h2, e2, groups[0].eps2, lists)) return -1;
    } else {
        // A block of at most kFusedMaxGroups i-groups with one softening
        // predicts j inside the force kernel (the predictor launch and the
        // JPred round trip dominate such blocks); any other block runs the
        // predictor first.
        const bool fused = E.opt.fused_small && ngroups == 1 && ni <= 64 * kFusedMaxGroups &&
                           choose_geometry(ni, nj).variant == 0;
        const int ld = stage_iblock(ni, xi, vi, excl, h2, e2);
        if (ld < 0) return -1;
        if (scatter_with_iblock(ld)) return -1;
        if (!fused && predict(nj, false)) return -1;
        const IBlock ib = iblock_at(E.d_ibuf, ld);
        // the reduction writes the results straight into pinned host memory;
        // for blocks of at most kDoneWordMaxI i the last group's reduction
        // also stores the completion word (larger blocks gained nothing: each
        // of their many reduction workgroups writes back its L2 first; with
        // kernel timing on, lasthalf waits for the event: it reads the timers)
        const Results r = results_at(E.h_res_dev, ni);
        const bool word = E.h_flag_dev && E.opt.done_word && !E.opt.timing && ni <= kDoneWordMaxI;
        const DoneWord dw{E.d_done_ticket, E.h_flag_dev, E.flag_seq + 1};
        for (int g = 0; g < ngroups; ++g) {
            // a group is the block shifted by its first slot (same stride); the
            // partial-sum buffer is reused in stream order
            const int o = groups[g].off;
            IBlock gb = ib;
            gb.xh += o; gb.yh += o; gb.zh += o; gb.xl += o; gb.yl += o; gb.zl += o;
            gb.vx += o; gb.vy += o; gb.vz += o; gb.excl += o;
            gb.x64 += o; gb.y64 += o; gb.z64 += o;
            gb.vx64 += o; gb.vy64 += o; gb.vz64 += o;
            Results gr = r;
            gr.acc += 3 * (size_t)o; gr.jerk += 3 * (size_t)o; gr.pot += o;
            gr.nn += o; gr.flags += o;
            const FusedPredict fp{E.d_jraw, E.ti, E.opt.predict_order, E.d_xp64, E.d_vp64,
                                  E.dcap};
            if (fused) {
                // xp64 / vp64 now hold this block's prediction, JPred an older one
                E.pred_valid = false;
                E.xp_valid = true;
                E.xp_ti = E.ti;
                E.xp_nj = nj;
            }
            if (enqueue_force(groups[g].n, nj, groups[g].eps2, gb, gr, groups[g].zskip,
                              fused ? &fp : nullptr, nullptr,
                              word && g == ngroups - 1 ? &dw : nullptr))
                return -1;
        }
        if (word) {
            E.flag_seq += 1;
            E.flag_pending = true;
        }
    }
    E.res_seq = mark_queued();
    // a block whose completion word lasthalf polls needs no event: if the
    // poll gives up, lasthalf waits for the stream instead
    if (!E.flag_pending) HIPCHK(hipEventRecord(E.ev_done, E.stream));
    E.done_seq = E.op_seq;
    E.last_ni = ni;
    E.last_nj = nj;
    E.last_include_self = include_self;
    E.calc_pending = true;
    E.calc_failed = false;
    return 0;
}

// Poll the block's completion word for seq, at most ~2 ms (a slower block,
// or a failed one, then waits for the stream instead).
bool poll_done(unsigned long long seq)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    for (unsigned k = 0;; ++k) {
        if (__atomic_load_n(E.h_flag, __ATOMIC_ACQUIRE) == seq) return true;
        if ((k & 255) == 255 && clk::now() - t0 > std::chrono::milliseconds(2)) return false;
        __builtin_ia32_pause();
    }
}

const Results *calc_wait(const char *who)
{
    static Results h;
    if (!E.calc_pending) {
        if (E.calc_failed) fail(1, "%s: the first half failed", who);
        else fail(1, "%s: no calculation in flight", who);
        return nullptr;
    }
    E.calc_pending = false;
    if (E.flag_pending) {
        E.flag_pending = false;
        if (poll_done(E.flag_seq)) {
            // the results are in host memory, and every host buffer the block
            // used is free: k_small reads its j updates and i-block from its
            // arguments, and on the FP32 path the scatter that read the pinned
            // staging ran before the reduction that stored the word; the
            // kernel's own end is ordered before the next launch
            E.synced_seq = std::max(E.synced_seq, E.done_seq);
            h = results_at(E.h_res, E.last_ni);
            E.last_refinements = count_refined(h, E.last_ni);
            return &h;
        }
        // no event was recorded for this block: the stream holds nothing after it
        const hipError_t e = hipStreamSynchronize(E.stream);
        if (e != hipSuccess) {
            E.calc_failed = true;
            fail(-1, "%s: %s", who, hipGetErrorString(e));
            return nullptr;
        }
        E.synced_seq = std::max(E.synced_seq, E.done_seq);
        h = results_at(E.h_res, E.last_ni);
        E.last_refinements = count_refined(h, E.last_ni);
        return &h;
    }
    hipError_t e = hipEventSynchronize(E.ev_done);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        E.calc_failed = true;
        fail(-1, "%s: %s", who, hipGetErrorString(e));
        return nullptr;
    }
    E.synced_seq = std::max(E.synced_seq, E.done_seq);
    harvest_timing();
    h = results_at(E.h_res, E.last_ni);
    E.last_refinements = count_refined(h, E.last_ni);
    return &h;
}

int ensure_nb(int ni)
{
    const size_t need = (size_t)ni * (size_t)E.opt.nb_cap;
    if (need > E.nbcap_entries) {
        dfree(E.d_nblist);
        E.nbcap_entries = 0;
        HIPCHK(hipMalloc((void **)&E.d_nblist, sizeof(int) * need));
        E.nbcap_entries = need;
    }
    if (ni > E.nbcount_cap) {
        hfree(E.h_nbcount);
        E.nbcount_cap = 0;
        HIPCHK(hipHostMalloc((void **)&E.h_nbcount, sizeof(int) * (size_t)ni,
                             hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer((void **)&E.h_nbcount_dev, E.h_nbcount, 0));
        E.nbcount_cap = ni;
    }
    return 0;
}

int read_nl()
{
    if (ensure_open("read_neighbour_list")) return -1;
    if (E.last_ni <= 0 || E.calc_failed) return fail(-1, "read_neighbour_list: no calculation");
    if (E.calc_pending) {
        // a block with a completion word recorded no event
        if (E.flag_pending) HIPCHK(hipStreamSynchronize(E.stream));
        else HIPCHK(hipEventSynchronize(E.ev_done));
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
    const size_t need = nl_scratch_ints(nslots, E.last_nj);
    if (need > E.nlscratch_cap) {
        dfree(E.d_nlscratch);
        E.nlscratch_cap = 0;
        HIPCHK(hipMalloc((void **)&E.d_nlscratch, sizeof(int) * need));
        E.nlscratch_cap = need;
    }
    if (nslots + 1 > E.nlinfo_cap) {
        hfree(E.h_nlinfo);
        E.nlinfo_cap = 0;
        const int c = std::max(nslots + 1, 1024);
        HIPCHK(hipHostMalloc((void **)&E.h_nlinfo, sizeof(int) * (size_t)c, hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer((void **)&E.h_nlinfo_dev, E.h_nlinfo, 0));
        E.nlinfo_cap = c;
    }
    if (E.h_nblist_cap == 0) {   // typical lists land here directly
        const size_t c = (size_t)1 << 20;
        HIPCHK(hipHostMalloc((void **)&E.h_nblist, sizeof(int) * c, hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer((void **)&E.h_nblist_dev, E.h_nblist, 0));
        E.h_nblist_cap = c;
    }
    // counts of unlisted slots stay 0; the kernels read slots and write
    // counts, list starts and (when they fit) the lists in pinned memory
    memset(E.h_nbcount, 0, sizeof(int) * (size_t)ni);
    {
        KTimer t(K_NBLIST);
        HIPCHK(launch_nblist(E.stream, E.d_xp64, E.dcap, E.d_jindex, E.last_nj, ib,
                             h2_at(E.d_ibuf, ld), e2_at(E.d_ibuf, ld), E.h_slots_dev, nslots,
                             E.last_include_self, E.opt.nb_cap, E.h_nbcount_dev, E.d_nlscratch,
                             E.d_nblist, E.h_nlinfo_dev, E.h_nblist_dev,
                             (int)std::min<size_t>(E.h_nblist_cap, INT_MAX)));
        t.stop();
    }
    HIPCHK(hipStreamSynchronize(E.stream));
    E.synced_seq = E.op_seq;
    harvest_timing();
    const size_t total = (size_t)E.h_nlinfo[nslots];
    int overflow = 0;
    for (int s = 0; s < nslots; ++s) {
        const int k = E.h_slots[s];
        if (E.h_nbcount[k] > E.opt.nb_cap) overflow = 1;
        E.nb_offset[k] = (size_t)E.h_nlinfo[s];
    }
    if (total > E.h_nblist_cap) {
        // too long for the pinned buffer: the kernel wrote them on the device
        hfree(E.h_nblist);
        E.h_nblist_cap = 0;
        HIPCHK(hipHostMalloc((void **)&E.h_nblist, sizeof(int) * total, hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer((void **)&E.h_nblist_dev, E.h_nblist, 0));
        E.h_nblist_cap = total;
        HIPCHK(hipMemcpyAsync(E.h_nblist, E.d_nblist, sizeof(int) * total,
                              hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
    }
    return overflow;
}

int nl_count(int slot)
{
    if (!E.nb_valid || slot < 0 || slot >= E.last_ni) return 0;
    return E.h_nbcount[slot];
}

const int *nl_list(int slot)
{
    if (!E.nb_valid || slot < 0 || slot >= E.last_ni) return nullptr;
    return E.h_nblist + E.nb_offset[slot];
}

int nl_max_count()
{
    int c = 0;
    if (E.nb_valid)
        for (int k = 0; k < E.last_ni; ++k) c = std::max(c, E.h_nbcount[k]);
    return c;
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
    ib.vx64 = E.d_vp64 + i_first;
    ib.vy64 = E.d_vp64 + (size_t)E.dcap + i_first;
    ib.vz64 = E.d_vp64 + 2 * (size_t)E.dcap + i_first;
    return 0;
}

// Workspace of self_sweep_into: the i-block's ten FP32/index arrays
// (ld * 40 B), the refinement flags (ld ints) and the j-split partial sums of
// the force geometry for (ni, nj), each start aligned to 256 bytes.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sweep_workspace_bytes(int nj, int ni)
{
    const int ld = (ni + 63) & ~63;
    const Geometry g = choose_geometry(ni, nj);
    return align256((size_t)ld * 40) + align256((size_t)ld * sizeof(int)) +
           partials_bytes(ld, g.npart);
}

int self_sweep_into(const char *who, int nj, int i_first, int ni, double ti, double eps2,
                    const Results &out, char *ws, size_t ws_bytes)
{
    if (ensure_open(who)) return -1;
    if (ni <= 0 || i_first < 0 || i_first + ni > nj || nj <= 0)
        return fail(-1, "%s: bad range i_first=%d ni=%d nj=%d", who, i_first, ni, nj);
    const size_t need = sweep_workspace_bytes(nj, ni);
    if (need > ws_bytes)
        return fail(-1, "%s: ni = %d, nj = %d needs %zu bytes of scratch, more than the %zu "
                        "it has", who, ni, nj, need, ws_bytes);
    if (E.calc_pending) HIPCHK(hipStreamSynchronize(E.stream));
    E.calc_pending = false;
    E.nb_valid = false;
    E.calc_failed = true;      // no neighbour lists for sweeps
    E.sweep_results = false;   // the results went to the caller's arrays
    // event timers would create events: off for this call
    const int timing = E.opt.timing;
    E.opt.timing = 0;
    E.ti = ti;
    int rc = flush_j(nj, false, true);
    if (rc == 0) rc = predict(nj, true);
    if (rc == 0) {
        const int ld = (ni + 63) & ~63;
        float *ibuf = reinterpret_cast<float *>(ws);
        hipError_t e = launch_iprep_self(E.stream, E.d_jpred, E.d_xp64, E.dcap, i_first, ni,
                                         ibuf, ld);
        if (e != hipSuccess) {
            rc = fail(-1, "%s: %s", who, hipGetErrorString(e));
        } else {
            IBlock ib = iblock_at(ws, ld);
            ib.x64 = E.d_xp64 + i_first;
            ib.y64 = E.d_xp64 + (size_t)E.dcap + i_first;
            ib.z64 = E.d_xp64 + 2 * (size_t)E.dcap + i_first;
            ib.vx64 = E.d_vp64 + i_first;
            ib.vy64 = E.d_vp64 + (size_t)E.dcap + i_first;
            ib.vz64 = E.d_vp64 + 2 * (size_t)E.dcap + i_first;
            char *flags = ws + align256((size_t)ld * 40);
            char *part = flags + align256((size_t)ld * sizeof(int));
            Results r = out;
            r.flags = reinterpret_cast<int *>(flags);
            rc = enqueue_force(ni, nj, eps2, ib, r, false, nullptr, part);
        }
    }
    E.opt.timing = timing;
    if (rc) return rc;
    E.last_ni = ni;
    E.last_nj = nj;
    return 0;
}

// j-split of a sweep with n_iw waves of i against nj j in tiles of `tile`:
// about waves_target waves in all (the built-in rule of choose_geometry).
void sweep_split(int n_iw, int nj, int tile, int &nsplit, int &jchunk)
{
    const int target = std::max(1, E.opt.waves_target);
    const int ntiles = std::max(1, (nj + tile - 1) / tile);
    int sp = E.opt.nsplit > 0 ? E.opt.nsplit : std::max(1, (target + n_iw - 1) / n_iw);
    sp = std::min(sp, ntiles);
    jchunk = ((ntiles + sp - 1) / sp) * tile;
    nsplit = std::max(1, (nj + jchunk - 1) / jchunk);
}

int sweep_begin(const char *who, int nj, int i_first, int ni, double ti)
{
    if (ensure_open(who)) return -1;
    if (ni <= 0 || i_first < 0 || i_first + ni > nj || nj <= 0)
        return fail(-1, "%s: bad range i_first=%d ni=%d nj=%d", who, i_first, ni, nj);
    if (E.calc_pending) HIPCHK(hipStreamSynchronize(E.stream));
    E.calc_pending = false;
    E.nb_valid = false;
    E.calc_failed = true;   // a sweep leaves no i-block for neighbour lists
    E.sweep_results = false;   // potential / densities reuse the results buffer
    E.ti = ti;
    if (flush_j(nj)) return -1;
    if (predict(nj, false)) return -1;
    return ensure_i(ni);
}

int potential_sweep(int nj, int i_first, int ni, double ti, double eps2, double *pot)
{
    if (sweep_begin("g6x_potential", nj, i_first, ni, ti)) return -1;
    IBlock ib;
    if (self_block(nj, i_first, ni, ib)) return -1;
    const int n_iw = (ni + 63) / 64;
    const int tile = (long long)n_iw * ((nj + kTile - 1) / kTile) >= E.opt.waves_target ? kTile : 64;
    int nsplit, jchunk;
    sweep_split(n_iw, nj, tile, nsplit, jchunk);
    if (ensure_partials((size_t)nsplit * ib.ld * sizeof(double))) return -1;
    const Results r = results_at(E.d_res, ni);
    {
        KTimer t(K_POT);
        HIPCHK(launch_pot(E.stream, E.d_jpred, nj, ib, ni, (float)eps2, nsplit, jchunk, tile,
                          reinterpret_cast<double *>(E.d_part), r.pot, E.opt.xcd_remap != 0));
        t.stop();
    }
    const Results h = results_at(E.h_res, ni);
    HIPCHK(hipMemcpyAsync(h.pot, r.pot, sizeof(double) * (size_t)ni, hipMemcpyDeviceToHost,
                          E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    harvest_timing();
    memcpy(pot, h.pot, sizeof(double) * (size_t)ni);
    E.last_ni = ni;
    E.last_nj = nj;
    return 0;
}

// waves of the k-nearest-neighbour sweep: 8 per CU (measured at N = 65536:
// 4.3 ms with 2 splits of 32768 j against 11.7 ms with 32 splits)
constexpr int kKnnWaves = 2048;

int density_sweep(int nj, int i_first, int ni, double ti, int k, double *density, double *dk2)
{
    if (k < 2 || k > 32) return fail(-1, "g6x_densities: k = %d outside [2, 32]", k);
    if (nj - 1 < k) return fail(-1, "g6x_densities: k = %d but only %d other particles", k, nj - 1);
    if (sweep_begin("g6x_densities", nj, i_first, ni, ti)) return -1;
    const int ld = (ni + 63) & ~63;
    const int n_iw = (ni + 63) / 64;
    // j-splits: each split refills its k-list from empty, and while the list
    // fills, some lane of the wave inserts at nearly every j; long splits pay
    // that once.  About kKnnWaves waves in all, splits of at least 256 j, at
    // most 64 of them ("nsplit" overrides).
    int nsplit = E.opt.nsplit > 0 ? E.opt.nsplit : std::max(1, (kKnnWaves + n_iw - 1) / n_iw);
    nsplit = std::min({nsplit, std::max(1, nj / 256), 64});
    const int jchunk = (nj + nsplit - 1) / nsplit;
    nsplit = (nj + jchunk - 1) / jchunk;
    const int kmax = knn_kmax(k);
    const size_t entries = (size_t)nsplit * kmax * ld;
    if (ensure_partials(entries * (sizeof(double) + sizeof(int)))) return -1;
    double *pd = reinterpret_cast<double *>(E.d_part);
    int *pa = reinterpret_cast<int *>(pd + entries);
    const Results r = results_at(E.d_res, ni);   // acc area: density[ni], dk2[ni]
    {
        KTimer t(K_KNN);
        HIPCHK(launch_knn(E.stream, E.d_xp64, E.dcap, nj, i_first, ni, k, nsplit, jchunk, pd, pa,
                          ld, E.d_jraw, r.acc, r.acc + ni));
        t.stop();
    }
    const Results h = results_at(E.h_res, ni);
    HIPCHK(hipMemcpyAsync(h.acc, r.acc, sizeof(double) * 2 * (size_t)ni, hipMemcpyDeviceToHost,
                          E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    harvest_timing();
    if (density) memcpy(density, h.acc, sizeof(double) * (size_t)ni);
    if (dk2) memcpy(dk2, h.acc + ni, sizeof(double) * (size_t)ni);
    E.last_ni = ni;
    E.last_nj = nj;
    return 0;
}

// Mean ms of force + reduction for one shape (after one untimed launch).
double time_shape(const Shape &sh, int nj, int ni, const IBlock &ib, double eps2, int reps)
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
    // the candidates overwrite the i-block and the results buffer: the last
    // block's neighbour lists and a self-sweep's results are gone
    E.last_ni = 0;
    E.calc_failed = true;
    E.nb_valid = false;
    E.sweep_results = false;
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
            for (int tile : {16, 32, 64, 128, kTile}) {
                if (variant == 1 && tile != kTile) continue;
                if (tile < 64 && ni > 64) continue;   // short tiles only for one i-group
                for (int w : waves) {
                    for (int wpg : {1, 4, 16}) {
                        if (variant == 1 && wpg != 1) continue;
                        Shape sh;
                        sh.variant = variant; sh.tile = tile; sh.waves = w; sh.lanes = 0;
                        sh.wpg = wpg;
                        const double ms = time_shape(sh, nj, ni, ib, eps2, reps);
                        if (ms < 0) { E.opt.timing = saved_timing; return fail(-1, "g6x_autotune: launch failed"); }
                        if (ms < best_ms) { best_ms = ms; best = sh; }
                    }
                }
            }
        }
        // neighbouring waves-per-workgroup counts for the winner
        if (best.variant == 0) {
            const Shape base = best;
            for (int wpg : {2, 8}) {
                Shape sh = base;
                sh.wpg = wpg;
                const double ms = time_shape(sh, nj, ni, ib, eps2, reps);
                if (ms >= 0 && ms < best_ms) { best_ms = ms; best = sh; }
            }
        }
        // reduction lanes for the winning force shape
        Shape lb = best;
        for (int lanes : {4, 8, 16, 32, 64}) {
            Shape sh = best;
            sh.lanes = lanes;
            const double ms = time_shape(sh, nj, ni, ib, eps2, reps);
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
    fprintf(f, "# %s force shapes: bucket (ni <= 2^bucket) variant tile waves lanes wpg\n",
            lib_tag());
    for (int b = 0; b < kBuckets; ++b) {
        const Shape &t = E.tuned[b];
        if (t.waves > 0)
            fprintf(f, "%d %d %d %d %d %d\n", b, t.variant, t.tile, t.waves, t.lanes, t.wpg);
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
        int b, v, t, w, l, q = 0;   // a table without the wpg column uses the rule
        const int got = sscanf(line, "%d %d %d %d %d %d", &b, &v, &t, &w, &l, &q);
        if (got != 5 && got != 6) continue;
        if (b < 0 || b >= kBuckets || (v != 0 && v != 1) || w <= 0 ||
            !valid_tile(t) || (v == 1 && t != kTile) ||
            (l != 0 && l != 4 && l != 8 && l != 16 && l != 32 && l != 64) ||
            (q != 0 && !valid_wpg(q)))
            continue;
        Shape &s = E.tuned[b];
        s.variant = v; s.tile = t; s.waves = w; s.lanes = l; s.wpg = q;
        ++n;
    }
    fclose(f);
    return n;
}


}  // namespace g6
