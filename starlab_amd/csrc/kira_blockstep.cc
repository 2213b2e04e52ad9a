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
        h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
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
            abi.firsthalf(&cl, &nj, &nii, ix, X, V, A, J, pt + s, &e2, h2.data());
            const int rc = abi.lasthalf2(&cl, &nj, &nii, ix, X, V, &e2, h2.data(), A, J, pt + s,
                                         nn + s);
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

}  // namespace

extern "C" {

const char *bs_last_error(void) { return g_err.c_str(); }

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
        // set_first_timestep (hdyn_ev.C:377-410)
        const double eta_init = eta / 16.0, cap = std::min(dt_max, t_end);
        for (int i = 0; i < n; ++i) {
            const double *ai = &R.a[3 * (size_t)i], *ji = &R.j[3 * (size_t)i];
            const double aa = dot(ai, ai), jj = dot(ji, ji);
            const double dt_adot = jj > 0 ? aa / jj : 1e300;
            const double dt_ff = (R.pot[i] < 0 && aa > 0) ? -R.pot[i] / aa : 1e300;
            double d = eta_init * std::sqrt(std::min(dt_adot, dt_ff));
            d = std::min(std::min(d, eta_init), cap);
            R.dt[i] = pow2_floor(d);
        }
        schedule(all);
        prev = all;
        while (!sched.empty()) {
            const double t_next = sched.begin()->first;
            if (t_next > t_end) break;
            blk.swap(sched.begin()->second);
            sched.erase(sched.begin());
            std::sort(blk.begin(), blk.end());
            // the previous block's particles go to j-memory (hdyn_grape6.C:3112-3113)
            if ((status = R.send_j(prev))) break;
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
            if ((status = R.force(t_next, blk, xi.data(), vi.data(), a1.data(), j1.data(),
                                  p1.data(), nn.data())))
                break;
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
            // get_derivs + corrector (hdyn_ev.C:1235-1250, 1717-1718), Aarseth
            // step (hdyn_ev.C:627-655), power-of-two rule (hdyn_ev.C:866-887)
            for (int q = 0; q < ni; ++q) {
                const int i = blk[q];
                const double d = t_next - R.t[i];
                double bt2[3], at3[3];
                for (int c = 0; c < 3; ++c) {
                    const size_t o = 3 * (size_t)i + c;
                    const double a0 = R.a[o], j0 = R.j[o];
                    const double aN = a1[3 * q + c], jN = j1[3 * q + c];
                    bt2[c] = -3 * (a0 - aN) - d * (2 * j0 + jN);
                    at3[c] = 2 * (a0 - aN) + d * (j0 + jN);
                }
                for (int c = 0; c < 3; ++c) {
                    const size_t o = 3 * (size_t)i + c;
                    R.x[o] = xi[3 * q + c] + (0.05 * at3[c] + kOne12 * bt2[c]) * d * d;
                    R.v[o] = vi[3 * q + c] + (0.25 * at3[c] + kOne3 * bt2[c]) * d;
                    R.a[o] = a1[3 * q + c];
                    R.j[o] = j1[3 * q + c];
                    R.k18[o] = bt2[c] / (9.0 * d * d);
                }
                R.pot[i] = p1[q];
                R.t[i] = t_next;
                const double dd2 = d * d, dd4 = dd2 * dd2, dd6 = dd4 * dd2;
                const double l2 = 36 * dot(at3, at3) / dd6;
                const double k2 = 4 * dot(bt2, bt2) / dd4;
                const double jj = dot(&j1[3 * q], &j1[3 * q]);
                const double aa = dot(&a1[3 * q], &a1[3 * q]);
                double nat = eta * std::sqrt((std::sqrt(aa * k2) + jj) / (std::sqrt(jj * l2) + k2));
                if (!(nat > 0)) nat = INFINITY;
                const double T = R.dt[i];
                double nd = T;
                if (nat < T)
                    nd = 0.5 * T;
                else if (nat >= 2 * T && std::fmod(t_next, 2 * T) == 0.0 && 2 * T <= dt_max)
                    nd = 2 * T;
                R.dt[i] = nd;
            }
            schedule(blk);
            prev.swap(blk);
            ++blocks;
            steps += ni;
            if (t_next == t_end) break;
        }
    } while (false);

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

}  // extern "C"
