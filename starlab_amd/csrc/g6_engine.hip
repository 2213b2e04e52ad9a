// g6_engine.hip -- the GRAPE-6 C ABI of libg6 (include/grape6.h) and its
// extensions (include/g6x.h), as a thin front end over the engine of
// g6_core.hip.
//
// The call sequence kira uses (hdyn_grape6.C:474-910, 1991-2181, 3067-3411)
// is open -> set_j* -> set_ti -> firsthalf -> lasthalf2 -> [read_nl ->
// get_nl*], repeated, with close/re-open and resets at any point.
//
// GRAPE-6 semantics kept here:
//   - self-exclusion by index: slot k skips the j whose index field equals
//     index[k] (hdyn_grape6.C:3462-3468), resolved to an address on the host;
//   - neighbour lists {j : d^2 < h2[k]} without softening, ascending address,
//     the i itself included when g6_npipes_() > 48 (hdyn_grape6.C:2309-2315).
#include "g6_core.h"
#include "../../include/grape6.h"
#include "../../include/g6x.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace g6 {
const char *lib_tag() { return "libg6"; }
const char *env_prefix() { return "G6_"; }

namespace {

// Scratch of g6x_self_sweep_noalloc: a fixed device array, so that call
// allocates nothing.  Each GPU the library runs on holds its own copy until
// the process ends.  Host code must not pass `scratch` itself to a kernel
// (on the host it only names the symbol): scratch_dev[device] is the
// device address, looked up once per device when the library opens it.
constexpr size_t kScratchBytes = (size_t)48 << 20;   // the default geometry needs <= 36.5 MiB for ni <= 65536
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];
char *scratch_dev[kMaxDevices] = {};

int setup_scratch()
{
    const int dev = E.device;
    if (dev < 0 || dev >= kMaxDevices) return 0;   // the entry point refuses on such a device
    if (scratch_dev[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(-1, "open: no device address for the static scratch on device %d (%s)", dev,
                    hipGetErrorString(e));
    scratch_dev[dev] = static_cast<char *>(p);
    return 0;
}

int open_device(int dev)
{
    const int rc = do_open(dev);
    if (rc) return rc;
    return setup_scratch();
}

int firsthalf(int nj, int ni, const int *index, const double (*xi)[3],
              const double (*vi)[3], double eps2, const double *h2)
{
    if (ni <= 0 || ni > E.opt.npipes) {
        E.calc_pending = false;
        E.calc_failed = true;
        return fail(-1, "g6calc_firsthalf: ni = %d outside [1, npipes = %d]", ni, E.opt.npipes);
    }
    if (!index || !xi || !vi) {
        E.calc_pending = false;
        E.calc_failed = true;
        return fail(-1, "g6calc_firsthalf: null index, position or velocity array");
    }
    std::vector<int> excl(ni);
    for (int k = 0; k < ni; ++k) excl[k] = excl_address(index[k]);
    const IGroup g{0, ni, eps2, false};
    return calc_begin("g6calc_firsthalf", nj, ni, xi, vi, excl.data(), h2, nullptr, &g, 1,
                      E.opt.npipes > 48 ? 1 : 0);
}

int lasthalf(int ni, double (*acc)[3], double (*jerk)[3], double *pot, int *nn)
{
    const Results *h = calc_wait("g6calc_lasthalf");
    if (!h) return 1;
    if (E.opt.fail_lasthalf > 0) {
        --E.opt.fail_lasthalf;
        return fail(1, "g6calc_lasthalf: injected failure (option fail_lasthalf)");
    }
    const int n = std::min(ni, E.last_ni);
    memcpy(acc, h->acc, sizeof(double) * 3 * (size_t)n);
    memcpy(jerk, h->jerk, sizeof(double) * 3 * (size_t)n);
    memcpy(pot, h->pot, sizeof(double) * (size_t)n);
    if (nn) memcpy(nn, h->nn, sizeof(int) * (size_t)n);
    if (nn && n > 0 && E.opt.fail_nn > 0) {
        --E.opt.fail_nn;
        nn[0] = -1;
    }
    return 0;
}

int read_nl_checked()
{
    if (E.opt.fail_read_nl > 0) {
        --E.opt.fail_read_nl;
        return fail(-1, "g6_read_neighbour_list: injected failure (option fail_read_nl)");
    }
    return read_nl();
}

}  // namespace
}  // namespace g6

using namespace g6;

#pragma GCC visibility push(default)

// ===========================================================================
// GRAPE-6 ABI
// ===========================================================================
extern "C" {

void g6_set_tunit_(int *) {}
void g6_set_xunit_(int *) {}

int g6_open_(int *)
{
    read_env();
    return open_device(E.opt.device);
}

int g6_close_(int *) { return do_close(); }

int g6_open_special(int ngpu, int *list)
{
    read_env();
    return open_device(ngpu > 0 && list ? list[0] : E.opt.device);
}

int g6_open_special2(int ngpu, int *list, int)
{
    return g6_open_special(ngpu, list);
}

int g6_reset_(int *)
{
    reset_state();
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

int g6_read_neighbour_list_(int *) { return read_nl_checked(); }
int g6_read_neighbour_list_old_(int *) { return read_nl_checked(); }

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
    if (n < 0) return fail(-1, "g6x_set_j_particles: n = %d", n);
    if (n > 0 && (!address || !index || !tj || !mass || !k18 || !j6 || !a2 || !v || !x))
        return fail(-1, "g6x_set_j_particles: null array (only dtj may be NULL)");
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
    E.sweep_results = false;   // until this sweep is queued
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
    if (enqueue_force(ni, nj, eps2, ib, r)) return -1;
    E.last_ni = ni;
    E.last_nj = nj;
    E.calc_failed = true;   // no neighbour lists for sweeps
    E.sweep_results = true;
    return 0;
}

int g6x_potential(int, int nj, int i_first, int ni, double ti, double eps2, double *pot)
{
    if (!pot) return fail(-1, "g6x_potential: pot is NULL");
    return potential_sweep(nj, i_first, ni, ti, eps2, pot);
}

int g6x_densities(int, int nj, int i_first, int ni, double ti, int k, double *density,
                  double *dk2)
{
    return density_sweep(nj, i_first, ni, ti, k, density, dk2);
}

int g6x_read_results(int, int ni, double *acc, double *jerk, double *pot, int *nn)
{
    if (ensure_open("g6x_read_results")) return -1;
    if (!E.sweep_results)
        return fail(-1, "g6x_read_results: the last calculation was not a g6x_self_sweep");
    if (ni < 0 || ni > E.last_ni)
        return fail(-1, "g6x_read_results: ni %d outside [0, %d]", ni, E.last_ni);
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
    if (nj < 0) return fail(-1, "g6x_read_predicted: nj = %d", nj);
    if (!E.xp_valid || nj > E.xp_nj) {
        // the last calculation did not keep an FP64 prediction (a small
        // block without neighbour spheres): predict at the current ti
        if (E.n_addr == 0 || nj > E.n_addr)
            return fail(-1, "g6x_read_predicted: no prediction");
        if (flush_j(nj) || predict(nj, true)) return -1;
    }
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

int g6x_self_sweep_noalloc(int, int nj, int i_first, int ni, double ti, double eps2,
                           double *acc, double *jerk, double *pot, int *nnbindex)
{
    const char *who = "g6x_self_sweep_noalloc";
    if (ensure_open(who)) return -1;
    if (!acc || !jerk || !pot || !nnbindex) return fail(-1, "%s: null output array", who);
    if (E.device < 0 || E.device >= kMaxDevices || !scratch_dev[E.device])
        return fail(-1, "%s: no static scratch on device %d", who, E.device);
    Results out{};
    out.acc = acc;
    out.jerk = jerk;
    out.pot = pot;
    out.nn = nnbindex;
    return self_sweep_into(who, nj, i_first, ni, ti, eps2, out, scratch_dev[E.device],
                           kScratchBytes);
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
    else if (k == "fused_small") o.fused_small = value != 0 ? 1 : 0;
    else if (k == "small_fp64") o.small_fp64 = value != 0 ? 1 : 0;
    else if (k == "done_word") o.done_word = value != 0 ? 1 : 0;
    else if (k == "fail_lasthalf") o.fail_lasthalf = (int)std::max(0LL, value);
    else if (k == "fail_read_nl") o.fail_read_nl = (int)std::max(0LL, value);
    else if (k == "fail_nn") o.fail_nn = (int)std::max(0LL, value);
    else if (k == "wpg") {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
            return -1;
        o.wpg = (int)value;
    }
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
    if (k == "fail_lasthalf") return o.fail_lasthalf;
    if (k == "fail_read_nl") return o.fail_read_nl;
    if (k == "fail_nn") return o.fail_nn;
    if (k == "fused_small") return o.fused_small;
    if (k == "small_fp64") return o.small_fp64;
    if (k == "done_word") return o.done_word;
    if (k == "last_tile") return E.geom_tile;
    if (k == "reduce_lanes") return o.reduce_lanes;
    if (k == "wpg") return o.wpg;
    if (k == "last_wpg") return E.geom_wpg;
    if (k == "last_lanes") return E.geom_lanes;
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
#pragma GCC visibility pop
