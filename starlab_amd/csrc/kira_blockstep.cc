// kira_blockstep.cc -- kira's block-step Hermite integration driving a
// GRAPE-6 ABI library, for top-level point masses (include/kira_blockstep.h).
//
// The same rules as starlab_amd/blockstep.py (BlockHermite), in the same
// operation order, compiled without floating-point contraction: the two give
// bitwise equal runs on the same library.  Citations are to the Starlab tree.
#include "kira_blockstep.h"
#include "grape6.h"
#include "g6x.h"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

namespace {

std::string g_err;

// Wall time of the last bs_run by phase (seconds): the j sends, firsthalf,
// lasthalf2, the host prediction of the block, correction and scheduling,
// and the whole run.
enum Phase { PH_SEND, PH_FIRST, PH_LAST, PH_PREDICT, PH_CORRECT, PH_TOTAL, PH_N };
double g_phase[PH_N];

int bs_fail(int code, const char *msg, const char *detail = "")
{
    g_err = std::string(msg) + detail;
    return code;
}

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr double kOne3 = 1.0 / 3.0;
constexpr double kOne12 = 1.0 / 12.0;

// Starlab's vec dot product: x*x + y*y + z*z, left to right
inline double dot(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// 2^floor(log2(d)) for finite d > 0, exactly
inline double pow2_floor(double d)
{
    int e = 0;
    std::frexp(d, &e);
    return std::ldexp(1.0, e - 1);
}

struct Abi {
    void *h = nullptr;
    decltype(&g6_open_) open = nullptr;
    decltype(&g6_close_) close = nullptr;
    decltype(&g6_npipes_) npipes = nullptr;
    decltype(&g6_set_ti_) set_ti = nullptr;
    decltype(&g6_set_j_particle_) set_j = nullptr;
    decltype(&g6calc_firsthalf_) firsthalf = nullptr;
    decltype(&g6calc_lasthalf2_) lasthalf2 = nullptr;
    decltype(&g6x_set_j_particles) set_j_bulk = nullptr;   // optional

    int load(const char *path)
    {
        // never unloaded: a GRAPE library may own threads (a GPU runtime, an
        // OpenMP pool) that would outlive its code; kira links it for good
        h = dlopen(path, RTLD_NOW | RTLD_LOCAL | RTLD_NODELETE);
        if (!h) return bs_fail(-1, "bs_run: dlopen failed: ", dlerror());
        const char *missing = nullptr;
        auto sym = [&](auto &f, const char *name) {
            f = reinterpret_cast<std::remove_reference_t<decltype(f)>>(dlsym(h, name));
            if (!f && !missing) missing = name;
        };
        sym(open, "g6_open_");
        sym(close, "g6_close_");
        sym(npipes, "g6_npipes_");
        sym(set_ti, "g6_set_ti_");
        sym(set_j, "g6_set_j_particle_");
        sym(firsthalf, "g6calc_firsthalf_");
        sym(lasthalf2, "g6calc_lasthalf2_");
        if (missing) return bs_fail(-1, "bs_run: the library does not export ", missing);
        set_j_bulk = reinterpret_cast<decltype(set_j_bulk)>(dlsym(h, "g6x_set_j_particles"));
        return 0;
    }
    ~Abi()
    {
        if (h) dlclose(h);
    }
};

struct Run {
    Abi abi;
    int n = 0;
    bool bulk = false;
    double eta = 0.1, eps2 = 0, dt_max = 0.125;
    std::vector<double> m, x, v, a, j, k18, pot, t, dt;
    int npipes = 1;
    double abi_s = 0;
    std::vector<double> sbuf;   // bulk j records

    // send_j_node_to_grape for particles idx (hdyn_grape6.C:197-288):
    // address = index = particle number, a/2, j/6, k/18 as stored
    int send_j(const std::vector<int> &idx)
    {
        const double t0 = now_s();
        int cl = 0;
        if (bulk && abi.set_j_bulk) {
            const int k = (int)idx.size();
            // tj, dtj, m (k each), then k18, j6, a2, v, x (3k each)
            sbuf.resize((size_t)k * 18);
            double *tj = sbuf.data(), *dtj = tj + k, *mm = dtj + k, *kk = mm + k;
            double *j6 = kk + 3 * (size_t)k, *a2 = j6 + 3 * (size_t)k;
            double *vv = a2 + 3 * (size_t)k, *xx = vv + 3 * (size_t)k;
            for (int q = 0; q < k; ++q) {
                const int i = idx[q];
                tj[q] = t[i];
                dtj[q] = dt[i];
                mm[q] = m[i];
                for (int c = 0; c < 3; ++c) {
                    kk[3 * q + c] = k18[3 * i + c];
                    j6[3 * q + c] = j[3 * i + c] / 6.0;
                    a2[3 * q + c] = a[3 * i + c] / 2.0;
                    vv[3 * q + c] = v[3 * i + c];
                    xx[3 * q + c] = x[3 * i + c];
                }
            }
            const int rc = abi.set_j_bulk(0, k, idx.data(), idx.data(), tj, dtj, mm, kk, j6, a2,
                                          vv, xx);
            abi_s += now_s() - t0;
            return rc ? bs_fail(-1, "bs_run: g6x_set_j_particles failed") : 0;
        }
        for (int i : idx) {
            int ad = i, ix = i;
            double tj = t[i], dtj = dt[i], mm = m[i];
            double kk[3], j6[3], a2[3], vv[3], xx[3];
            for (int c = 0; c < 3; ++c) {
                kk[c] = k18[3 * i + c];
                j6[c] = j[3 * i + c] / 6.0;
                a2[c] = a[3 * i + c] / 2.0;
                vv[c] = v[3 * i + c];
                xx[c] = x[3 * i + c];
            }
            if (abi.set_j(&cl, &ad, &ix, &tj, &dtj, &mm, kk, j6, a2, vv, xx)) {
                abi_s += now_s() - t0;
                return bs_fail(-1, "bs_run: g6_set_j_particle_ failed");
            }
        }
        abi_s += now_s() - t0;
        return 0;
    }

    // g6_set_ti_ + firsthalf/lasthalf2 in chunks of npipes (hdyn_grape6.C:3219-3338)
    int force(double ti, const std::vector<int> &idx, const double *xi, const double *vi,
              double *acc, double *jrk, double *pt, int *nn)
    {
        const double t0 = now_s();
        int cl = 0;
        double tt = ti;
        abi.set_ti(&cl, &tt);
        const int ni = (int)idx.size();
        std::vector<double> h2((size_t)npipes, 0.0);
        for (int s = 0; s < ni; s += npipes) {
            int nii = std::min(npipes, ni - s), nj = n;
            double e2 = eps2;
            int *ix = const_cast<int *>(idx.data()) + s;
            auto X = reinterpret_cast<double(*)[3]>(const_cast<double *>(xi) + 3 * (size_t)s);
            auto V = reinterpret_cast<double(*)[3]>(const_cast<double *>(vi) + 3 * (size_t)s);
            auto A = reinterpret_cast<double(*)[3]>(acc + 3 * (size_t)s);
            auto J = reinterpret_cast<double(*)[3]>(jrk + 3 * (size_t)s);
            const double tf = now_s();
            abi.firsthalf(&cl, &nj, &nii, ix, X, V, A, J, pt + s, &e2, h2.data());
            const double tl = now_s();
            const int rc = abi.lasthalf2(&cl, &nj, &nii, ix, X, V, &e2, h2.data(), A, J, pt + s,
                                         nn + s);
            g_phase[PH_FIRST] += tl - tf;
            g_phase[PH_LAST] += now_s() - tl;
            if (rc) {
                abi_s += now_s() - t0;
                char buf[64];
                snprintf(buf, sizeof buf, "%d", rc);
                return bs_fail(-1, "bs_run: g6calc_lasthalf2_ returned ", buf);
            }
        }
        abi_s += now_s() - t0;
        return 0;
    }
};


// set_first_timestep (hdyn_ev.C:377-410) for every particle
void first_timesteps(Run &R, double t_end)
{
    const double eta_init = R.eta / 16.0, cap = std::min(R.dt_max, t_end);
    for (int i = 0; i < R.n; ++i) {
        const double *ai = &R.a[3 * (size_t)i], *ji = &R.j[3 * (size_t)i];
        const double aa = dot(ai, ai), jj = dot(ji, ji);
        const double dt_adot = jj > 0 ? aa / jj : 1e300;
        const double dt_ff = (R.pot[i] < 0 && aa > 0) ? -R.pot[i] / aa : 1e300;
        double d = eta_init * std::sqrt(std::min(dt_adot, dt_ff));
        d = std::min(std::min(d, eta_init), cap);
        R.dt[i] = pow2_floor(d);
    }
}

// get_derivs + corrector (hdyn_ev.C:1235-1250, 1717-1718), Aarseth step
// (hdyn_ev.C:627-655), power-of-two rule (hdyn_ev.C:866-887) for particle i
// predicted to t_next (xi, vi) with new acc a1, jerk j1 and pot p1
void correct_one(Run &R, int i, double t_next, const double *xi, const double *vi,
                 const double *a1, const double *j1, double p1)
{
    const double d = t_next - R.t[i];
    double bt2[3], at3[3];
    for (int c = 0; c < 3; ++c) {
        const size_t o = 3 * (size_t)i + c;
        const double a0 = R.a[o], j0 = R.j[o];
        const double aN = a1[c], jN = j1[c];
        bt2[c] = -3 * (a0 - aN) - d * (2 * j0 + jN);
        at3[c] = 2 * (a0 - aN) + d * (j0 + jN);
    }
    for (int c = 0; c < 3; ++c) {
        const size_t o = 3 * (size_t)i + c;
        R.x[o] = xi[c] + (0.05 * at3[c] + kOne12 * bt2[c]) * d * d;
        R.v[o] = vi[c] + (0.25 * at3[c] + kOne3 * bt2[c]) * d;
        R.a[o] = a1[c];
        R.j[o] = j1[c];
        R.k18[o] = bt2[c] / (9.0 * d * d);
    }
    R.pot[i] = p1;
    R.t[i] = t_next;
    const double dd2 = d * d, dd4 = dd2 * dd2, dd6 = dd4 * dd2;
    const double l2 = 36 * dot(at3, at3) / dd6;
    const double k2 = 4 * dot(bt2, bt2) / dd4;
    const double jj = dot(j1, j1);
    const double aa = dot(a1, a1);
    double nat = R.eta * std::sqrt((std::sqrt(aa * k2) + jj) / (std::sqrt(jj * l2) + k2));
    if (!(nat > 0)) nat = INFINITY;
    const double T = R.dt[i];
    double nd = T;
    if (nat < T)
        nd = 0.5 * T;
    else if (nat >= 2 * T && std::fmod(t_next, 2 * T) == 0.0 && 2 * T <= R.dt_max)
        nd = 2 * T;
    R.dt[i] = nd;
}

// ---------------------------------------------------------------------------
// bs_run_kira: kira's GRAPE-6 call mix (hdyn_grape6.C) around the same
// integrator.  Node k is top-level node k: GRAPE index = address = k.
// ---------------------------------------------------------------------------
constexpr int kNRetry = 5;                            // hdyn_grape6.C:75
constexpr int kMaxPerturbers = 256;                   // hdyn.h:25
constexpr int kMaxNeighbors = 16 * kMaxPerturbers;    // hdyn_grape6.C:2262
constexpr double kRnbIncrease = 2.0;                  // :2263
constexpr double kRnbDecrease = 0.8;                  // :2264
constexpr int kMaxForceCount = 20;                    // :2801
constexpr int kNDens = 12;                            // :3604
constexpr double kVeryLarge = 1.0e300;                // VERY_LARGE_NUMBER, stdinc.h:206
constexpr double kDMinFac = 0.25;                     // DEFAULT_D_MIN_FAC, kira_defaults.h:9

struct KiraAbi {
    Abi base;
    decltype(&g6_reset_) reset = nullptr;
    decltype(&g6_reset_fofpga_) reset_fofpga = nullptr;
    decltype(&g6_read_neighbour_list_) read_nl = nullptr;
    decltype(&g6_get_neighbour_list_) get_nl = nullptr;

    int load(const char *path)
    {
        if (base.load(path)) return -1;
        reset = reinterpret_cast<decltype(reset)>(dlsym(base.h, "g6_reset_"));
        reset_fofpga = reinterpret_cast<decltype(reset_fofpga)>(dlsym(base.h, "g6_reset_fofpga_"));
        read_nl = reinterpret_cast<decltype(read_nl)>(dlsym(base.h, "g6_read_neighbour_list_"));
        get_nl = reinterpret_cast<decltype(get_nl)>(dlsym(base.h, "g6_get_neighbour_list_"));
        if (!reset || !reset_fofpga || !read_nl || !get_nl)
            return bs_fail(-1, "bs_run_kira: the library lacks the reset or neighbour-list calls");
        return 0;
    }
};

enum Stat {
    S_FORCE = 4, S_READ, S_READ_OVF, S_GET, S_GET_OVF, S_REDO, S_PERTURB, S_ENERGY, S_DENSITY,
    S_RESET, S_REOPEN, S_RELOAD, S_NN_LIST, S_PERT_OVF, S_ERR42, S_ELOG, S_COUNT
};

struct Kira {
    KiraAbi abi;
    Run R;                       // integrator state (R.abi unused: calls go through abi)
    bs_kira_opts o{};
    bs_call_fn obs = nullptr;
    void *user = nullptr;
    int n = 0, npipes = 1;
    double st[S_COUNT] = {0};
    double gamma23 = 0, d_min_sq = 0, h2_crit = 0;
    int coll_freq = 8, pert_freq = 1;

    // j-memory as sent (per address)
    std::vector<double> jt, jdt, jm, jx, jv, ja2, jj6, jk18;
    // per node
    std::vector<double> radius, rnb, d_nn_sq, d_coll_sq;
    std::vector<int> nn, coll, nb_count, npert, valid_pert, bin_of;
    std::vector<double> px, pv, pt;    // host prediction cache (pt = its time)
    std::vector<double> na, nja, np;   // results of the last force call per node
    std::vector<int> nnn;
    std::vector<double> density;

    bool is_open = false, potential_used = false, attached_once = false;
    double offset = 0;               // grape_time_offset
    double sys_t = 0;                // time of the block in progress
    std::vector<int> prev;           // previous_nodes (sent at the next block)

    // i arrays of one call (n_pipes long, as kira passes them)
    std::vector<int> iindex, inn;
    std::vector<double> ipos, ivel, iacc, ijerk, ipot, ih2;
    std::vector<int> nbl;
    std::string stop;

    bool is_parent(int k) const { return bin_of[k] >= 0; }

    int observe(const bs_call &c)
    {
        if (!obs) return 0;
        const bs_jmem_sent jmem{n, jt.data(), jdt.data(), jm.data(), jx.data(), jv.data(),
                                ja2.data(), jj6.data(), jk18.data()};
        if (obs(user, &jmem, &c)) {
            stop = "bs_run_kira: stopped by the observer";
            return -2;
        }
        return 0;
    }

    // host prediction of node k to time t (predict_loworder, _dyn_.h:290-295)
    void predict(int k, double t)
    {
        if (pt[k] == t) return;
        const double d = t - R.t[k], d3 = d * kOne3, d2 = d * 0.5;
        for (int c = 0; c < 3; ++c) {
            const size_t o = 3 * (size_t)k + c;
            px[o] = ((R.j[o] * d3 + R.a[o]) * d2 + R.v[o]) * d + R.x[o];
            pv[o] = (R.j[o] * d2 + R.a[o]) * d + R.v[o];
        }
        pt[k] = t;
    }

    // send_j_node_to_grape (hdyn_grape6.C:197-288)
    int send_j(int k, bool energy)
    {
        const double t0 = now_s();
        int cl = 0, ad = k, ix = k;
        double tj = energy ? 0.0 : R.t[k] - offset;
        double dtj = R.dt[k] > 0 ? R.dt[k] : 1.0;
        double mm = R.m[k];
        double kk[3], j6[3], a2[3], vv[3], xx[3];
        for (int c = 0; c < 3; ++c) {
            const size_t o = 3 * (size_t)k + c;
            kk[c] = R.k18[o];
            j6[c] = R.j[o] / 6.0;
            a2[c] = R.a[o] / 2.0;
            vv[c] = R.v[o];
            xx[c] = R.x[o];
        }
        const int rc = abi.base.set_j(&cl, &ad, &ix, &tj, &dtj, &mm, kk, j6, a2, vv, xx);
        R.abi_s += now_s() - t0;
        if (rc) return bs_fail(-1, "bs_run_kira: g6_set_j_particle_ failed");
        jt[k] = tj; jdt[k] = dtj; jm[k] = mm;
        for (int c = 0; c < 3; ++c) {
            const size_t o = 3 * (size_t)k + c;
            jx[o] = xx[c]; jv[o] = vv[c]; ja2[o] = a2[c]; jj6[o] = j6[c]; jk18[o] = kk[c];
        }
        return 0;
    }

    // initialize_grape_arrays (hdyn_grape6.C:292-376)
    int initialize(bool reset_counters)
    {
        while (sys_t - offset > 1) offset += 1;
        for (int k = 0; k < n; ++k) {
            if (send_j(k, false)) return -1;
            if (reset_counters) nb_count[k] = 0;
        }
        st[S_RELOAD] += 1;
        return 0;
    }

    int open_device()
    {
        int cl = 0;
        const double t0 = now_s();
        const int rc = abi.base.open(&cl);
        R.abi_s += now_s() - t0;
        if (rc) return bs_fail(-1, "bs_run_kira: g6_open_ failed");
        is_open = true;
        return 0;
    }

    // an open by the driver itself (not a reset): the first attaches, later
    // ones re-attach after a release
    int attach()
    {
        if (open_device()) return -1;
        if (attached_once) st[S_REOPEN] += 1;
        attached_once = true;
        return 0;
    }

    // reset_grape (hdyn_grape6.C:405-438)
    int reset_grape()
    {
        int cl = 0;
        const double t0 = now_s();
        abi.reset(&cl);
        abi.reset_fofpga(&cl);
        abi.base.close(&cl);
        R.abi_s += now_s() - t0;
        is_open = false;
        st[S_RESET] += 1;
        if (open_device()) return -1;
        return initialize(false);
    }

    // check_reattach_grape6 (hdyn_grape6.C:3004-3054)
    int check_reattach()
    {
        bool restart = false, reattached = false;
        if (!is_open) {
            if (attach()) return -1;
            reattached = true;
            restart = true;
        }
        if (potential_used) {
            restart = true;
            potential_used = false;
        }
        if (restart) {
            if (initialize(!reattached)) return -1;
            prev.clear();
        }
        return 0;
    }

    // force_by_grape (hdyn_grape6.C:474-910): one chunk of at most n_pipes
    // nodes; returns 0, or the error that persisted through kNRetry resets
    // (< 0: a failure of the run itself)
    int force_by_grape(int kind, const int *nodes, int ni, bool pot_only, int level = 0)
    {
        if (ni <= 0) return 0;
        int cl = 0, nj = n;
        double ti = pot_only ? 0.0 : sys_t - offset;
        const double t0 = now_s();
        abi.base.set_ti(&cl, &ti);
        for (int q = 0; q < ni; ++q) {
            const int k = nodes[q];
            iindex[q] = k;
            predict(k, sys_t);
            for (int c = 0; c < 3; ++c) {
                const size_t o = 3 * (size_t)k + c;
                ipos[3 * q + c] = pot_only ? R.x[o] : px[o];
                ivel[3 * q + c] = pv[o];
                iacc[3 * q + c] = R.a[o];
                ijerk[3 * q + c] = R.j[o];
            }
            ipot[q] = R.pot[k];
            ih2[q] = rnb[k];
        }
        for (int q = ni; q < npipes; ++q) {   // padding: copies of the last, h2 = 0
            iindex[q] = iindex[ni - 1];
            for (int c = 0; c < 3; ++c) {
                ipos[3 * q + c] = ipos[3 * (ni - 1) + c];
                ivel[3 * q + c] = ivel[3 * (ni - 1) + c];
                iacc[3 * q + c] = iacc[3 * (ni - 1) + c];
                ijerk[3 * q + c] = ijerk[3 * (ni - 1) + c];
            }
            ipot[q] = ipot[ni - 1];
            ih2[q] = 0;
        }
        double e2 = R.eps2;
        auto X = reinterpret_cast<double(*)[3]>(ipos.data());
        auto V = reinterpret_cast<double(*)[3]>(ivel.data());
        auto A = reinterpret_cast<double(*)[3]>(iacc.data());
        auto J = reinterpret_cast<double(*)[3]>(ijerk.data());
        abi.base.firsthalf(&cl, &nj, &ni, iindex.data(), X, V, A, J, ipot.data(), &e2, ih2.data());
        int error = abi.base.lasthalf2(&cl, &nj, &ni, iindex.data(), X, V, &e2, ih2.data(), A, J,
                                       ipot.data(), inn.data());
        R.abi_s += now_s() - t0;
        st[S_FORCE] += 1;
        {
            const bs_call c{kind, ti, nj, ni, iindex.data(), ipos.data(), ivel.data(), ih2.data(),
                            error, iacc.data(), ijerk.data(), ipot.data(), inn.data(),
                            BS_NOT_CALLED, 0, 0, 0, nullptr};
            if (observe(c)) return -2;
        }
        if (!error) {
            for (int q = 0; q < ni; ++q) {
                if (inn[q] < 0 || ipot[q] > 0) {
                    error = 42;
                    st[S_ERR42] += 1;
                    break;
                }
                if (inn[q] >= nj) inn[q] = 0;
            }
        }
        if (!error) {
            for (int q = 0; q < ni; ++q) {
                const int k = nodes[q];
                np[k] = ipot[q];
                if (pot_only) continue;
                for (int c = 0; c < 3; ++c) {
                    na[3 * (size_t)k + c] = iacc[3 * q + c];
                    nja[3 * (size_t)k + c] = ijerk[3 * q + c];
                }
                const int b = inn[q];
                nn[k] = b;
                predict(b, sys_t);
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = ipos[3 * q + c] - px[3 * (size_t)b + c];
                d_nn_sq[k] = dot(d, d);
            }
            return 0;
        }
        if (level < kNRetry) {
            if (reset_grape()) return -1;
            return force_by_grape(kind, nodes, ni, pot_only, level + 1);
        }
        return error;
    }

    // get_force_and_neighbors (hdyn_grape6.C:1991-2181): 0, 1 (list
    // storage overflow), 2 (a persistent error), < 0 (run failure)
    int force_and_neighbors(int kind, const int *nodes, int ni, bool need, int level = 0)
    {
        const int e = force_by_grape(kind, nodes, ni, false);
        if (e < 0) return e;
        if (e) return 2;
        if (!need) return 0;
        int cl = 0;
        const double t0 = now_s();
        const int status = abi.read_nl(&cl);
        R.abi_s += now_s() - t0;
        st[S_READ] += 1;
        {
            const bs_call c{kind, sys_t - offset, n, 0, nullptr, nullptr, nullptr, nullptr, 0,
                            nullptr, nullptr, nullptr, nullptr, status, 0, 0, 0, nullptr};
            if (observe(c)) return -2;
        }
        if (status > 0) {
            st[S_READ_OVF] += 1;
            return 1;
        }
        if (status < 0) {
            if (level < kNRetry) {
                if (reset_grape()) return -1;
                return force_and_neighbors(kind, nodes, ni, need, level + 1);
            }
            return 2;
        }
        return 0;
    }

    // binary_scale (hdyn_inline.C:150-190): max(semi-major axis, separation)
    double binary_scale(int k) const
    {
        const int b = bin_of[k];
        const double *dx = o.bin_dx + 3 * (size_t)b, *dv = o.bin_dv + 3 * (size_t)b;
        const double sep = std::sqrt(dot(dx, dx));
        const double energy = 0.5 * dot(dv, dv) - R.m[k] / sep;
        const double sma = 0.5 * R.m[k] / std::fabs(energy);
        return std::max(sma, sep);
    }

    // set_grape_neighbor_radius (hdyn_grape6.C:1851-1985)
    bool neighbor_radius(int k)
    {
        rnb[k] = 0;
        const bool nn_ok = nn[k] >= 0 && nn[k] != k && d_nn_sq[k] < 0.1 * kVeryLarge;
        if (!is_parent(k)) {
            if (nb_count[k] == 0 && radius[k] > 0) {
                double r = 4 * radius[k] * radius[k];
                if (nn_ok) {
                    r = std::max(r, d_nn_sq[k]);
                } else {
                    const double r90_sq = d_min_sq / (kDMinFac * kDMinFac);
                    r = std::max(r, r90_sq * std::pow((double)n, 4.0 / 3));
                }
                rnb[k] = r;
            }
        } else if (nb_count[k] == 0 || !valid_pert[k]) {
            npert[k] = 0;                       // new_perturber_list
            const double rb = binary_scale(k);
            double r = gamma23 * rb * rb;       // perturbation_scale_sq (hdyn_inline.C:241-254)
            if (r > 100) r = 100;
            if (nn_ok) r = std::max(r, d_nn_sq[k]);
            rnb[k] = r;
        }
        return rnb[k] > 0;
    }

    // sort_nodes_and_reduce_rnb (hdyn_grape6.C:2192-2255)
    int sort_and_reduce(int *list, int ni)
    {
        int jnext = ni, jmax = ni;
        double rmax = 0;
        for (int j = ni - 1; j >= 0; --j) {
            if (rnb[list[j]] > 0) {
                if (j < --jnext) std::swap(list[j], list[jnext]);
                if (rnb[list[jnext]] > rmax) {
                    jmax = jnext;
                    rmax = rnb[list[jnext]];
                }
            }
        }
        if (jmax != jnext) std::swap(list[jnext], list[jmax]);
        for (int j = jnext; j < ni; ++j)
            if (!is_parent(list[j]) || j == jnext) rnb[list[j]] *= 0.5;
        return jnext;
    }

    // one g6_get_neighbour_list_ call, reported to the observer
    int get_list(int k, int pipe, int &count)
    {
        int cl = 0, p = pipe, mx = kMaxNeighbors, len = 0;
        const double t0 = now_s();
        const int rc = abi.get_nl(&cl, &p, &mx, &len, nbl.data());
        R.abi_s += now_s() - t0;
        st[S_GET] += 1;
        if (rc) st[S_GET_OVF] += 1;
        count = len;
        predict(k, sys_t);
        const bs_call c{BS_CALL_LIST, sys_t - offset, n, 1, &k, &px[3 * (size_t)k], &pv[3 * (size_t)k],
                        &rnb[k], rc, nullptr, nullptr, nullptr, nullptr, BS_NOT_CALLED, pipe, mx,
                        std::min(len, mx), nbl.data()};
        if (observe(c)) return -2;
        return rc;
    }

    // is_perturber, hybrid criterion (hdyn_inline.C:289-324, kira_defaults.h:25)
    bool is_perturber(int k, double m_pert, double d2, double rpfac) const
    {
        const double d3 = d2 * std::sqrt(d2);
        if (m_pert <= 0.5 * R.m[k]) return d3 < rpfac;
        return 0.5 * R.m[k] * d3 < m_pert * rpfac;
    }

    // get_neighbors_and_adjust_h2 (hdyn_grape6.C:2277-2797): 0, 1 (list
    // overflow), 2 (nn or coll not found), < 0 (run failure)
    int neighbors_and_adjust(int k, int pipe)
    {
        int len = 0;
        const int rc = get_list(k, pipe, len);
        if (rc == -2) return -2;
        const int n_extra = npipes > 48 ? 1 : 0;
        if (rc) {
            if (len >= kMaxNeighbors) {
                rnb[k] *= 0.25;
                return 1;
            }
        }
        int status = 2;
        if (len > n_extra) {
            status = 0;
            double dmin = kVeryLarge, dcmin = kVeryLarge;
            int bmin = -1, cmin = -1, npl = 0;
            const bool want = is_parent(k) && !valid_pert[k];
            double rpfac = 0;
            if (want) {
                const double rb = binary_scale(k), rp2 = gamma23 * rb * rb;
                rpfac = rp2 * std::sqrt(rp2);   // define_perturbation_radius_factor, criterion 2
            }
            predict(k, sys_t);
            for (int q = 0; q < std::min(len, kMaxNeighbors); ++q) {
                const int b = nbl[q];
                if (b < 0 || b >= n || b == k) continue;
                predict(b, sys_t);
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = px[3 * (size_t)k + c] - px[3 * (size_t)b + c];
                const double d2 = dot(d, d);
                if (d2 < dmin) { dmin = d2; bmin = b; }          // update_nn_coll
                const double sr = radius[k] + radius[b];
                const double cd = d2 - sr * sr;
                if (cd < dcmin) { dcmin = cd; cmin = b; }
                if (want && is_perturber(k, R.m[b], d2, rpfac)) ++npl;
            }
            if (want) {
                npert[k] = npl;
                valid_pert[k] = npl <= kMaxPerturbers;
                if (npl > kMaxPerturbers) st[S_PERT_OVF] += 1;
            }
            if (bmin >= 0) {
                nn[k] = bmin;
                d_nn_sq[k] = dmin;
                st[S_NN_LIST] += 1;
            } else {
                status = 2;
            }
            if (cmin >= 0) {
                coll[k] = cmin;
                d_coll_sq[k] = dcmin;
            } else {
                status = 2;
            }
        } else if (is_parent(k)) {
            npert[k] = 0;
            valid_pert[k] = 1;
            if (nn[k] >= 0) {                  // EXPAND_TO_FIND_COLL
                d_coll_sq[k] = d_nn_sq[k];
                coll[k] = nn[k];
            } else {
                nn[k] = k;
            }
        }
        if (status) rnb[k] *= kRnbIncrease;
        return status;
    }

    // get_coll_and_perturbers (hdyn_grape6.C:2803-2998): nodes done
    int coll_and_perturbers(int *list, int ni)
    {
        int inext = 0;
        for (int ip = 0; ip < ni; ++ip) {
            int pipe = ip;
            const int k = list[ip];
            if (rnb[k] > 0) {
                int count_force = 1, status;
                while ((status = neighbors_and_adjust(k, pipe))) {
                    if (status < 0) return status;
                    if (status == 2 && is_parent(k)) break;
                    if (count_force > kMaxForceCount || (status == 2 && rnb[k] > h2_crit)) {
                        nn[k] = k;
                        d_nn_sq[k] = 2 * h2_crit;
                        break;
                    }
                    pipe = 0;
                    ip = ni;                   // the rest of the chunk is redone
                    int s;
                    while ((s = force_and_neighbors(BS_CALL_REDO, &k, 1, true))) {
                        if (s < 0) return s;
                        if (s == 2) return bs_fail(-1, "bs_run_kira: persistent error in a redo");
                        st[S_REDO] += 1;
                        rnb[k] *= kRnbDecrease;
                        ++count_force;
                    }
                    st[S_REDO] += 1;
                }
            }
            ++inext;
        }
        return inext;
    }

    // grape6_calculate_acc_and_jerk (hdyn_grape6.C:3067-3411) for the
    // block's nodes at sys_t; results in na / nja / np per node
    int calc_acc_and_jerk(std::vector<int> &cur)
    {
        if (check_reattach()) return -1;
        for (int k : prev)
            if (send_j(k, false)) return -1;
        const int n_top = (int)cur.size();
        prev = cur;
        int last_nbr = n_top;
        for (int i = n_top - 1; i >= 0; --i) {
            const int k = cur[i];
            const bool need = neighbor_radius(k);
            if (need && is_parent(k)) valid_pert[k] = 0;
            if (need && --last_nbr > i) std::swap(cur[last_nbr], cur[i]);   // SORT_NODE_LIST
        }
        int i = 0;
        while (i < n_top) {
            const int ni = std::min(npipes, n_top - i);
            bool need = false;
            for (int q = 0; q < ni; ++q)
                if (rnb[cur[i + q]] > 0) { need = true; break; }
            const int s = force_and_neighbors(BS_CALL_BLOCK, cur.data() + i, ni, need);
            if (s < 0) return s;
            if (s == 2) return bs_fail(-1, "bs_run_kira: persistent error in a block");
            if (s == 1) {
                i += sort_and_reduce(cur.data() + i, ni);
            } else if (need) {
                const int d = coll_and_perturbers(cur.data() + i, ni);
                if (d < 0) return d;
                i += d;
            } else {
                i += ni;
            }
        }
        for (int k : cur) {
            if (!is_parent(k)) {
                nb_count[k] = (nb_count[k] + 1) % coll_freq;
                if (coll[k] < 0) coll[k] = nn[k];
            } else {
                nb_count[k] = (nb_count[k] + 1) % pert_freq;
            }
        }
        return 0;
    }

    // grape6_calculate_perturbation (hdyn_grape6.C:3415-3587)
    int perturbation(int k)
    {
        if (check_reattach()) return -1;
        const int b = bin_of[k];
        const double m1 = o.bin_m1[b], m2 = R.m[k] - m1, M = R.m[k];
        const double *dx = o.bin_dx + 3 * (size_t)b, *dv = o.bin_dv + 3 * (size_t)b;
        predict(k, sys_t);
        int cl = 0, nj = n, ni = 2;
        double ti = sys_t - offset;
        for (int q = 0; q < 2; ++q) {
            const double f = q == 0 ? -m2 / M : m1 / M;
            iindex[q] = k;                     // both exclude the CM node
            for (int c = 0; c < 3; ++c) {
                ipos[3 * q + c] = px[3 * (size_t)k + c] + f * dx[c];
                ivel[3 * q + c] = pv[3 * (size_t)k + c] + f * dv[c];
            }
            const double mo = q == 0 ? m2 : m1;   // the other component
            const double r2 = dot(dx, dx), r = std::sqrt(r2), jj = mo / (r * r2);
            ipot[q] = jj * r2;
            for (int c = 0; c < 3; ++c) {
                iacc[3 * q + c] = jj * r;
                ijerk[3 * q + c] = jj;
            }
            ih2[q] = 0;
        }
        for (int q = 2; q < npipes; ++q) {
            iindex[q] = iindex[1];
            for (int c = 0; c < 3; ++c) {
                ipos[3 * q + c] = ipos[3 + c];
                ivel[3 * q + c] = ivel[3 + c];
                iacc[3 * q + c] = iacc[3 + c];
                ijerk[3 * q + c] = ijerk[3 + c];
            }
            ipot[q] = ipot[1];
            ih2[q] = 0;
        }
        double e2 = R.eps2;
        auto X = reinterpret_cast<double(*)[3]>(ipos.data());
        auto V = reinterpret_cast<double(*)[3]>(ivel.data());
        auto A = reinterpret_cast<double(*)[3]>(iacc.data());
        auto J = reinterpret_cast<double(*)[3]>(ijerk.data());
        const double t0 = now_s();
        abi.base.set_ti(&cl, &ti);
        abi.base.firsthalf(&cl, &nj, &ni, iindex.data(), X, V, A, J, ipot.data(), &e2, ih2.data());
        const int error = abi.base.lasthalf2(&cl, &nj, &ni, iindex.data(), X, V, &e2, ih2.data(), A,
                                             J, ipot.data(), inn.data());
        R.abi_s += now_s() - t0;
        st[S_FORCE] += 1;
        st[S_PERTURB] += 1;
        const bs_call c{BS_CALL_PERTURB, ti, nj, ni, iindex.data(), ipos.data(), ivel.data(),
                        ih2.data(), error, iacc.data(), ijerk.data(), ipot.data(), inn.data(),
                        BS_NOT_CALLED, 0, 0, 0, nullptr};
        if (observe(c)) return -2;
        return 0;                              // kira does not retry it (:3578-3586)
    }

    // grape6_calculate_energies, CM form (hdyn_grape6.C:1403-1546): every
    // node sent with t_j = 0, pot-only chunks at t_i = 0
    int energies(double &etot)
    {
        if (!is_open && attach()) return -1;
        std::vector<int> all(n);
        for (int k = 0; k < n; ++k) {
            all[k] = k;
            if (send_j(k, true)) return -1;
        }
        for (int i = 0; i < n; i += npipes) {
            const int e = force_by_grape(BS_CALL_ENERGY, all.data() + i, std::min(npipes, n - i), true);
            if (e < 0) return e;
            if (e) return bs_fail(-1, "bs_run_kira: persistent error in the energy sweep");
        }
        double epot = 0, ekin = 0;
        for (int k = 0; k < n; ++k) {
            const double *vk = &R.v[3 * (size_t)k];
            epot += 0.5 * R.m[k] * np[k];
            ekin += 0.5 * R.m[k] * dot(vk, vk);
        }
        etot = ekin + epot;
        potential_used = true;
        st[S_ENERGY] += 1;
        return 0;
    }

    // compute_density from a list (compute_density.C:69-226): k-th nearest
    // distance, masses of the k-1 nearer over (4/3) pi d_k^3
    double list_density(int k, int len)
    {
        std::vector<std::pair<double, int>> d;
        predict(k, sys_t);
        for (int q = 0; q < len; ++q) {
            const int b = nbl[q];
            if (b < 0 || b >= n || b == k) continue;
            predict(b, sys_t);
            double e[3];
            for (int c = 0; c < 3; ++c) e[c] = px[3 * (size_t)b + c] - px[3 * (size_t)k + c];
            d.push_back({dot(e, e), b});
        }
        std::stable_sort(d.begin(), d.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                             return a.first < b.first;
                         });
        if ((int)d.size() < kNDens) return 0;
        double mass = 0;
        for (int q = 0; q < kNDens - 1; ++q) mass += R.m[d[q].second];
        const double dk = std::sqrt(d[kNDens - 1].first);
        return dk > 0 ? mass / (4.0 / 3.0 * M_PI * dk * dk * dk) : 0;
    }

    // count_neighbors_and_adjust_h2 (hdyn_grape6.C:3704-3898)
    int count_and_adjust(int k, int pipe)
    {
        int len = 0;
        const int rc = get_list(k, pipe, len);
        if (rc == -2) return -2;
        const int n_extra = npipes > 48 ? 1 : 0;
        if (rc && len >= kMaxNeighbors) {
            rnb[k] *= std::pow(10.0 * kNDens / (double)len, 0.6666667);
            return -1;
        }
        if (rnb[k] <= 0) return 0;
        if (len < kNDens + n_extra) {
            const double fac = std::min(2.0, std::pow((kNDens + 3.0) / (1.0 + len - n_extra), 1.5));
            rnb[k] *= fac;
            return 1;
        }
        density[k] = list_density(k, std::min(len, kMaxNeighbors));
        rnb[k] = 0;
        return 0;
    }

    // get_densities (hdyn_grape6.C:4079-4280): 0, 1 (list storage
    // overflow), 2 (persistent error), -2 (observer stop)
    int get_densities(int *list, int ni, double h2c)
    {
        int n_force = 0;
        bool force = true;
        while (force) {
            if (++n_force > 20) {
                for (int ip = 0; ip < ni; ++ip)
                    if (rnb[list[ip]] > 0) {
                        rnb[list[ip]] = 0;
                        density[list[ip]] = 0;
                    }
                break;
            }
            const int s = force_and_neighbors(BS_CALL_DENSITY, list, ni, true);
            if (s < 0) return s;
            if (s) return s;
            force = false;
            for (int ip = 0; ip < ni; ++ip) {
                const int k = list[ip];
                if (rnb[k] > 0) {
                    const int c = count_and_adjust(k, ip);
                    if (c == -2) return -2;
                    if (c) {
                        if (rnb[k] > h2c) {
                            density[k] = 0;
                            rnb[k] = 0;
                        } else {
                            force = true;
                        }
                    }
                }
            }
        }
        return 0;
    }

    // grape6_calculate_densities (hdyn_grape6.C:4283-4449), h2_crit = 4
    int densities()
    {
        if (!is_open && attach()) return -1;
        if (initialize(true)) return -1;
        std::vector<int> top(n);
        const double r90_sq = d_min_sq / (kDMinFac * kDMinFac);
        const double rnn0 = r90_sq * std::pow((double)n, 4.0 / 3);
        for (int k = 0; k < n; ++k) {
            top[k] = k;
            double r = rnn0;                  // set_grape_density_radius (:3607-3685)
            if (nn[k] >= 0 && nn[k] != k && d_nn_sq[k] < 0.1 * kVeryLarge && d_nn_sq[k] > 0)
                r = d_nn_sq[k] > 100 * r ? 0 : 5 * d_nn_sq[k];
            else
                r = d_nn_sq[k] > 100 * r ? 0 : 5 * std::max(r, d_nn_sq[k]);
            rnb[k] = r;
            density[k] = 0;
        }
        int i = 0, count = 0;
        while (i < n) {
            const int ni = std::min(npipes, n - i);
            const int s = get_densities(top.data() + i, ni, 4.0);
            if (s < 0) return s;
            if (s == 2) return bs_fail(-1, "bs_run_kira: persistent error in a density round");
            if (s) {
                for (int j = i; j < i + ni; ++j) rnb[top[j]] *= 0.9;
                if (++count > 20) return bs_fail(-1, "bs_run_kira: density lists keep overflowing");
            } else {
                i += ni;
                count = 0;
            }
        }
        for (int k = 0; k < n; ++k) rnb[k] = 0;
        potential_used = true;
        st[S_DENSITY] += 1;
        return 0;
    }

    int log_point(int &n_log)
    {
        double e = 0;
        if (energies(e)) return -1;
        if (o.e_log && n_log < o.e_log_cap) o.e_log[n_log] = e;
        ++n_log;
        st[S_ELOG] = n_log;
        if (o.densities && densities()) return -1;
        return 0;
    }
};
}  // namespace

extern "C" {

const char *bs_last_error(void) { return g_err.c_str(); }

int bs_timers(double *seconds)
{
    if (!seconds) return bs_fail(-1, "bs_timers: null array");
    for (int k = 0; k < PH_N; ++k) seconds[k] = g_phase[k];
    return 0;
}

int bs_run(const char *grape_lib, int n, const double *m, double *x, double *v, double t_end,
           double eta, double eps2, double dt_max, int bulk, double *acc, double *jerk,
           double *pot, double *dt, double *stats, bs_block_fn observer, void *user)
{
    g_err.clear();
    if (n < 1 || !m || !x || !v || !grape_lib) return bs_fail(-1, "bs_run: bad arguments");
    // zero or non-finite steps would never reach t_end
    if (!(t_end > 0) || !(eta > 0) || !(eps2 >= 0) || !(dt_max > 0) || !std::isfinite(t_end) ||
        !std::isfinite(eta) || !std::isfinite(eps2) || !std::isfinite(dt_max))
        return bs_fail(-1, "bs_run: need t_end > 0, eta > 0, eps2 >= 0, dt_max > 0 (finite)");
    const double t_start = now_s();
    for (double &p : g_phase) p = 0;
    Run R;
    if (R.abi.load(grape_lib)) return -1;
    R.n = n;
    R.bulk = bulk != 0;
    R.eta = eta;
    R.eps2 = eps2;
    R.dt_max = dt_max;
    R.m.assign(m, m + n);
    R.x.assign(x, x + 3 * (size_t)n);
    R.v.assign(v, v + 3 * (size_t)n);
    R.a.assign(3 * (size_t)n, 0.0);
    R.j.assign(3 * (size_t)n, 0.0);
    R.k18.assign(3 * (size_t)n, 0.0);
    R.pot.assign(n, 0.0);
    R.t.assign(n, 0.0);
    R.dt.assign(n, 1.0);   // dtj of the initial load
    int cl = 0;
    if (R.abi.open(&cl)) return bs_fail(-1, "bs_run: g6_open_ failed");
    R.npipes = std::max(1, R.abi.npipes());

    long long blocks = 0, steps = 0;
    int status = 0;
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    std::vector<double> xi, vi, a1, j1, p1;
    std::vector<int> nn;
    std::map<double, std::vector<int>> sched;
    std::vector<int> prev, blk;

    auto schedule = [&](const std::vector<int> &idx) {
        for (int i : idx) sched[R.t[i] + R.dt[i]].push_back(i);
    };

    do {
        // initial forces at t = 0 on the unpredicted state, then the first steps
        if ((status = R.send_j(all))) break;
        a1.assign(3 * (size_t)n, 0.0); j1.assign(3 * (size_t)n, 0.0);
        p1.assign(n, 0.0); nn.assign(n, 0);
        if ((status = R.force(0.0, all, R.x.data(), R.v.data(), a1.data(), j1.data(), p1.data(),
                              nn.data())))
            break;
        R.a = a1;
        R.j = j1;
        R.pot = p1;
        first_timesteps(R, t_end);
        schedule(all);
        prev = all;
        while (!sched.empty()) {
            const double t_next = sched.begin()->first;
            if (t_next > t_end) break;
            blk.swap(sched.begin()->second);
            sched.erase(sched.begin());
            std::sort(blk.begin(), blk.end());
            // the previous block's particles go to j-memory (hdyn_grape6.C:3112-3113)
            double tp = now_s();
            if ((status = R.send_j(prev))) break;
            g_phase[PH_SEND] += now_s() - tp;
            tp = now_s();
            const int ni = (int)blk.size();
            xi.resize(3 * (size_t)ni); vi.resize(3 * (size_t)ni);
            a1.resize(3 * (size_t)ni); j1.resize(3 * (size_t)ni);
            p1.resize(ni); nn.resize(ni);
            // predict_loworder (include/_dyn_.h:290-295)
            for (int q = 0; q < ni; ++q) {
                const int i = blk[q];
                const double d = t_next - R.t[i], d3 = d * kOne3, d2 = d * 0.5;
                for (int c = 0; c < 3; ++c) {
                    const size_t o = 3 * (size_t)i + c;
                    xi[3 * q + c] = ((R.j[o] * d3 + R.a[o]) * d2 + R.v[o]) * d + R.x[o];
                    vi[3 * q + c] = (R.j[o] * d2 + R.a[o]) * d + R.v[o];
                }
            }
            g_phase[PH_PREDICT] += now_s() - tp;
            if ((status = R.force(t_next, blk, xi.data(), vi.data(), a1.data(), j1.data(),
                                  p1.data(), nn.data())))
                break;
            tp = now_s();
            if (observer) {
                // j-memory now equals the current state: every particle was last
                // sent as it is (the previous block just now)
                const bs_jmem jm{n, R.t.data(), R.m.data(), R.x.data(), R.v.data(),
                                 R.a.data(), R.j.data(), R.k18.data()};
                if (observer(user, &jm, t_next, ni, blk.data(), xi.data(), vi.data(),
                             a1.data(), j1.data(), p1.data(), nn.data())) {
                    status = -2;
                    g_err = "bs_run: stopped by the observer";
                    break;
                }
            }
            for (int q = 0; q < ni; ++q)
                correct_one(R, blk[q], t_next, &xi[3 * q], &vi[3 * q], &a1[3 * q], &j1[3 * q],
                            p1[q]);
            schedule(blk);
            prev.swap(blk);
            ++blocks;
            steps += ni;
            g_phase[PH_CORRECT] += now_s() - tp;
            if (t_next == t_end) break;
        }
    } while (false);
    g_phase[PH_TOTAL] = now_s() - t_start;

    R.abi.close(&cl);
    if (status) return status;
    std::memcpy(x, R.x.data(), sizeof(double) * 3 * (size_t)n);
    std::memcpy(v, R.v.data(), sizeof(double) * 3 * (size_t)n);
    if (acc) std::memcpy(acc, R.a.data(), sizeof(double) * 3 * (size_t)n);
    if (jerk) std::memcpy(jerk, R.j.data(), sizeof(double) * 3 * (size_t)n);
    if (pot) std::memcpy(pot, R.pot.data(), sizeof(double) * (size_t)n);
    if (dt) std::memcpy(dt, R.dt.data(), sizeof(double) * (size_t)n);
    if (stats) {
        stats[0] = (double)blocks;
        stats[1] = (double)steps;
        stats[2] = R.abi_s;
        stats[3] = now_s() - t_start;
    }
    return 0;
}

int bs_run_kira(const char *grape_lib, int n, const double *m, double *x, double *v, double t_end,
                double eta, double eps2, double dt_max, const bs_kira_opts *opts, double *acc,
                double *jerk, double *pot, double *dt, double *stats, bs_call_fn observer,
                void *user)
{
    g_err.clear();
    if (n < 2 || !m || !x || !v || !grape_lib || !opts) return bs_fail(-1, "bs_run_kira: bad arguments");
    if (!(t_end > 0) || !(eta > 0) || !(eps2 >= 0) || !(dt_max > 0) || !std::isfinite(t_end) ||
        !std::isfinite(eta) || !std::isfinite(eps2) || !std::isfinite(dt_max))
        return bs_fail(-1, "bs_run_kira: need t_end > 0, eta > 0, eps2 >= 0, dt_max > 0 (finite)");
    const bs_kira_opts &o = *opts;
    if (o.nbin < 0 || (o.nbin > 0 && (!o.bin_node || !o.bin_m1 || !o.bin_dx || !o.bin_dv)) ||
        !(o.star_radius >= 0) || !(o.dt_log >= 0) || o.release_every < 0)
        return bs_fail(-1, "bs_run_kira: bad options");
    const double t_start = now_s();
    Kira K;
    if (K.abi.load(grape_lib)) return -1;
    K.o = o;
    K.obs = observer;
    K.user = user;
    K.n = n;
    Run &R = K.R;
    R.n = n;
    R.eta = eta;
    R.eps2 = eps2;
    R.dt_max = dt_max;
    R.m.assign(m, m + n);
    R.x.assign(x, x + 3 * (size_t)n);
    R.v.assign(v, v + 3 * (size_t)n);
    R.a.assign(3 * (size_t)n, 0.0);
    R.j.assign(3 * (size_t)n, 0.0);
    R.k18.assign(3 * (size_t)n, 0.0);
    R.pot.assign(n, 0.0);
    R.t.assign(n, 0.0);
    R.dt.assign(n, 1.0);
    K.coll_freq = o.coll_freq > 0 ? o.coll_freq : 8;
    K.pert_freq = o.pert_freq > 0 ? o.pert_freq : 1;
    K.gamma23 = std::pow(o.gamma > 0 ? o.gamma : 1e-7, -2.0 / 3);
    K.d_min_sq = (kDMinFac / n) * (kDMinFac / n);   // kira_init.C:212-222
    K.h2_crit = 8192 * K.d_min_sq;                   // hdyn_grape6.C:3187
    for (auto *vec : {&K.jt, &K.jdt, &K.jm, &K.radius, &K.rnb, &K.np, &K.density, &K.pt})
        vec->assign(n, 0.0);
    for (auto *vec : {&K.jx, &K.jv, &K.ja2, &K.jj6, &K.jk18, &K.px, &K.pv, &K.na, &K.nja})
        vec->assign(3 * (size_t)n, 0.0);
    K.pt.assign(n, NAN);
    K.d_nn_sq.assign(n, kVeryLarge);
    K.d_coll_sq.assign(n, kVeryLarge);
    K.nn.assign(n, -1);
    K.coll.assign(n, -1);
    K.nb_count.assign(n, 0);
    K.npert.assign(n, 0);
    K.valid_pert.assign(n, 0);
    K.bin_of.assign(n, -1);
    for (int b = 0; b < o.nbin; ++b) {
        const int k = o.bin_node[b];
        if (k < 0 || k >= n || K.bin_of[k] >= 0 || !(o.bin_m1[b] > 0) || !(o.bin_m1[b] < m[k]))
            return bs_fail(-1, "bs_run_kira: bad binary node or component mass");
        K.bin_of[k] = b;
    }
    for (int k = 0; k < n; ++k)
        if (K.bin_of[k] < 0) K.radius[k] = o.star_radius;
    // the library's npipes sizes the i arrays (kira: hdyn_grape6.C:445-468)
    K.npipes = std::max(2, K.abi.base.npipes());
    K.iindex.assign(K.npipes, 0);
    K.inn.assign(K.npipes, 0);
    K.ipot.assign(K.npipes, 0.0);
    K.ih2.assign(K.npipes, 0.0);
    for (auto *vec : {&K.ipos, &K.ivel, &K.iacc, &K.ijerk}) vec->assign(3 * (size_t)K.npipes, 0.0);
    K.nbl.assign(kMaxNeighbors, 0);

    long long blocks = 0, steps = 0;
    int status = 0, n_log = 0;
    double next_log = o.dt_log > 0 ? 0.0 : INFINITY;
    std::map<double, std::vector<int>> sched;
    std::vector<int> blk, cur;
    std::vector<double> xi, vi, a1, j1;
    auto schedule = [&](const std::vector<int> &idx) {
        for (int i : idx) sched[R.t[i] + R.dt[i]].push_back(i);
    };
    auto fail_status = [&](int rc) {
        if (rc == -2) g_err = K.stop;
        return rc == -2 ? -2 : -1;
    };

    do {
        // initial forces: every node at t = 0 (the first call attaches and
        // loads j-memory), then the first steps
        K.sys_t = 0;
        cur.resize(n);
        for (int k = 0; k < n; ++k) cur[k] = k;
        if ((status = K.calc_acc_and_jerk(cur))) { status = fail_status(status); break; }
        for (int k = 0; k < n; ++k) {
            for (int c = 0; c < 3; ++c) {
                R.a[3 * (size_t)k + c] = K.na[3 * (size_t)k + c];
                R.j[3 * (size_t)k + c] = K.nja[3 * (size_t)k + c];
            }
            R.pot[k] = K.np[k];
        }
        K.pt.assign(n, NAN);
        first_timesteps(R, t_end);
        std::vector<int> all(n);
        for (int k = 0; k < n; ++k) all[k] = k;
        schedule(all);
        if (next_log == 0.0) {
            if ((status = K.log_point(n_log))) { status = fail_status(status); break; }
            next_log += o.dt_log;
        }
        while (!sched.empty()) {
            const double t_next = sched.begin()->first;
            if (t_next > t_end) break;
            blk.swap(sched.begin()->second);
            sched.erase(sched.begin());
            std::sort(blk.begin(), blk.end());
            K.sys_t = t_next;
            cur = blk;
            if ((status = K.calc_acc_and_jerk(cur))) { status = fail_status(status); break; }
            for (int k : blk)
                if (K.is_parent(k) && (status = K.perturbation(k))) break;
            if (status) { status = fail_status(status); break; }
            const int ni = (int)blk.size();
            xi.resize(3 * (size_t)ni); vi.resize(3 * (size_t)ni);
            a1.resize(3 * (size_t)ni); j1.resize(3 * (size_t)ni);
            for (int q = 0; q < ni; ++q) {
                const int k = blk[q];
                K.predict(k, t_next);
                for (int c = 0; c < 3; ++c) {
                    xi[3 * q + c] = K.px[3 * (size_t)k + c];
                    vi[3 * q + c] = K.pv[3 * (size_t)k + c];
                    a1[3 * q + c] = K.na[3 * (size_t)k + c];
                    j1[3 * q + c] = K.nja[3 * (size_t)k + c];
                }
            }
            for (int q = 0; q < ni; ++q) {
                correct_one(R, blk[q], t_next, &xi[3 * q], &vi[3 * q], &a1[3 * q], &j1[3 * q],
                            K.np[blk[q]]);
                K.pt[blk[q]] = NAN;               // its state changed: predict again
            }
            schedule(blk);
            ++blocks;
            steps += ni;
            if (o.release_every > 0 && blocks % o.release_every == 0 && K.is_open) {
                int cl = 0;                       // check_release_grape6 (:958-987)
                const double t0 = now_s();
                K.abi.base.close(&cl);
                R.abi_s += now_s() - t0;
                K.is_open = false;
            }
            if (t_next >= next_log) {
                if ((status = K.log_point(n_log))) { status = fail_status(status); break; }
                while (next_log <= t_next) next_log += o.dt_log;
            }
            if (t_next == t_end) break;
        }
    } while (false);

    if (K.is_open) {
        int cl = 0;
        K.abi.base.close(&cl);
    }
    if (status) return status;
    std::memcpy(x, R.x.data(), sizeof(double) * 3 * (size_t)n);
    std::memcpy(v, R.v.data(), sizeof(double) * 3 * (size_t)n);
    if (acc) std::memcpy(acc, R.a.data(), sizeof(double) * 3 * (size_t)n);
    if (jerk) std::memcpy(jerk, R.j.data(), sizeof(double) * 3 * (size_t)n);
    if (pot) std::memcpy(pot, R.pot.data(), sizeof(double) * (size_t)n);
    if (dt) std::memcpy(dt, R.dt.data(), sizeof(double) * (size_t)n);
    if (o.density) std::memcpy(o.density, K.density.data(), sizeof(double) * (size_t)n);
    if (stats) {
        K.st[0] = (double)blocks;
        K.st[1] = (double)steps;
        K.st[2] = R.abi_s;
        K.st[3] = now_s() - t_start;
        for (int k = 0; k < S_COUNT; ++k) stats[k] = K.st[k];
    }
    return 0;
}

}  // extern "C"

