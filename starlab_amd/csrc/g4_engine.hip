// g4_engine.hip -- the GRAPE-4 (HARP-3) C ABI of libg4x (include/grape4.h),
// a front end over the engine of g6_core.hip: the same j-memory, predictor,
// force, reduction and neighbour-list kernels as libg6.
//
// kira's GRAPE-4 driver (src/node/dyn/hdyn/evolve/hdyn_grape4.C) calls
//   h3open_ -> h3mjpdma_indirect_/h3mjpdma_start_* (j loads, :240-295)
//   -> h3setti_ -> h3calc_firsthalf_ -> h3wait_ -> h3calc_lasthalf_ (:347-391)
//   -> h3nbread_ -> h3nblist_* (:419-420, :867, :1513)
// per i-block.  configure probes libg4x for h3open_ only when no GRAPE-6
// library is found, and then defines USE_HALF_CHIP (configure.ac:214-221): kira
// puts each i-particle in two adjacent pipes and pads the block to
// h3npipe_() slots with copies of the last particle (hdyn_grape4.C:353-384).
//
// Arithmetic semantics follow the reference's IEEE GRAPE-4 simulator
// (src/harpsims/harp3fastsim.c):
//   - j-memory address = hostindex - 1; a DMA buffer holds a snapshot of the
//     caller's arrays, written to j-memory by h3mjpdma_start_ (:140-201);
//   - predictor x + s(v + s(a2 + s j6)), v + 2s(a2 + 1.5 s j6), s = ti - tj
//     (:251-265), aj = a/2 and jj = j/6 as kira stores them (hdyn_grape4.C:263-264);
//   - eps2 and h2 are doubled internally (:240-241): r^2 = 2 eps2 + d^2;
//   - a pair with r^2 = 0 (the i itself when eps2 = 0) adds nothing (:304);
//   - neighbour list of pipe i = {j : r^2 < 2 h2[i]} in ascending address,
//     the i itself included (:300-303), read back per pipe by h3nblist_;
//   - pot is not computed by the simulator; here pot = -sum m_j / r over the
//     same pairs, the sign of kira's own potential (hdyn_ev.C:2031).
// Pair arithmetic is the libg6 law: FP32 on double-single positions with
// FP64 accumulation.
//
// Work done once per distinct i: consecutive pipes with bit-identical
// (x, v, eps2, h2) -- USE_HALF_CHIP pairs and the padding -- are computed
// once and their results copied, so a block of n particles costs n, not
// h3npipe_(), i-particles.  Pipes are grouped by softening for the kernels.
//
// Errors: the ABI's calls return nothing, so a failing call prints
// "libg4x: <reason>" and exits the process with status 1 (the reference
// simulator likewise exits on out-of-range sizes, harp3fastsim.c:155-158,
// 233-236); the option "exit_on_error" = 0 (env G4_EXIT_ON_ERROR=0) keeps the
// process alive and counts the failures instead (g4x_errors()).
#include "g6_core.h"
#include "../../include/grape4.h"
#include "../../include/g4x.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace g6 {
const char *lib_tag() { return "libg4x"; }
const char *env_prefix() { return "G4_"; }

namespace {

constexpr int kMaxDmaBuffers = 1024;

struct DmaRecord {
    int addr;
    double x[3], v[3], a2[3], j6[3], m, tj;
};

struct G4State {
    bool env_read = false;
    int jpmax = 4096;
    int exit_on_error = 1;
    int dedupe = 1;
    long long errors = 0;
    std::vector<std::vector<DmaRecord>> bufs;   // DMA buffers by id
    int buff_used = 0;                          // buffers taken with buff_id < 0
    // last i-block: pipe -> staged (distinct) slot
    std::vector<int> slot_of_pipe;
    int n_pipes = 0, n_slots = 0;
    // staging of the distinct i-particles, grouped by softening
    std::vector<double> xs, vs, h2s, e2s;
    std::vector<IGroup> groups;
};

G4State G;

void g4_read_env()
{
    if (G.env_read) return;
    G.env_read = true;
    read_env();
    auto env = [](const char *name, long long d) {
        const char *s = getenv(name);
        return (s && *s) ? atoll(s) : d;
    };
    G.jpmax = (int)std::max(1LL, env("G4_JPMAX", G.jpmax));
    G.exit_on_error = (int)env("G4_EXIT_ON_ERROR", G.exit_on_error);
    G.dedupe = (int)env("G4_DEDUPE", G.dedupe);
}

// A void entry point failed: the message is already printed by fail().
void fatal()
{
    ++G.errors;
    if (G.exit_on_error) {
        fprintf(stderr, "libg4x: exiting (set G4_EXIT_ON_ERROR=0 to continue)\n");
        fflush(stderr);
        exit(1);
    }
}

void dma_copy(int nj, const int *hostindex, const double (*xj)[3], const double (*vj)[3],
              const double (*aj)[3], const double (*jj)[3], const double *mj,
              const double *tj, int bufno)
{
    if (nj < 0) { fail(-1, "h3mjpdma_indirect: nj = %d", nj); fatal(); return; }
    if (nj > 0 && (!hostindex || !xj || !vj || !aj || !jj || !mj || !tj)) {
        fail(-1, "h3mjpdma_indirect: null array");
        fatal();
        return;
    }
    if (bufno < 0 || bufno >= kMaxDmaBuffers) {
        fail(-1, "h3mjpdma_indirect: buffer %d outside [0, %d)", bufno, kMaxDmaBuffers);
        fatal();
        return;
    }
    if ((int)G.bufs.size() <= bufno) G.bufs.resize(bufno + 1);
    std::vector<DmaRecord> &b = G.bufs[bufno];
    b.clear();
    b.reserve(nj);
    for (int k = 0; k < nj; ++k) {
        const int h = hostindex[k] - 1;
        if (h < 0) {
            fail(-1, "h3mjpdma_indirect: hostindex[%d] = %d (indices are 1-based)", k, h + 1);
            fatal();
            return;
        }
        DmaRecord r;
        r.addr = h;
        for (int c = 0; c < 3; ++c) {
            r.x[c] = xj[h][c]; r.v[c] = vj[h][c]; r.a2[c] = aj[h][c]; r.j6[c] = jj[h][c];
        }
        r.m = mj[h];
        r.tj = tj[h];
        b.push_back(r);
    }
}

void dma_start(int bufno)
{
    if (bufno < 0 || bufno >= (int)G.bufs.size()) return;   // an unused buffer holds nothing
    static const double zero[3] = {0.0, 0.0, 0.0};
    for (const DmaRecord &r : G.bufs[bufno]) {
        if (set_j(r.addr, r.addr, r.tj, 1.0, r.m, zero, r.j6, r.a2, r.v, r.x)) {
            fatal();
            return;
        }
    }
}

bool same_pipe(const double (*xi)[3], const double (*vi)[3], const double *eps2,
               const double *h2, int a, int b)
{
    return memcmp(xi[a], xi[b], sizeof(double) * 3) == 0 &&
           memcmp(vi[a], vi[b], sizeof(double) * 3) == 0 &&
           memcmp(&eps2[a], &eps2[b], sizeof(double)) == 0 &&
           memcmp(&h2[a], &h2[b], sizeof(double)) == 0;
}

int firsthalf(int nj, int ni, const double (*xi)[3], const double (*vi)[3],
              const double *eps2, const double *h2)
{
    g4_read_env();
    if (ni <= 0 || ni > E.opt.npipes) {
        E.calc_pending = false;
        E.calc_failed = true;
        return fail(-1, "h3calc_firsthalf: ni = %d outside [1, npipes = %d]", ni, E.opt.npipes);
    }
    if (!xi || !vi || !eps2 || !h2) {
        E.calc_pending = false;
        E.calc_failed = true;
        return fail(-1, "h3calc_firsthalf: null position, velocity, eps2 or h2 array");
    }
    // distinct i-particles in pipe order
    std::vector<int> first_pipe;       // distinct slot -> its first pipe
    G.slot_of_pipe.assign(ni, 0);
    for (int k = 0; k < ni; ++k) {
        if (G.dedupe && k > 0 && same_pipe(xi, vi, eps2, h2, k, k - 1)) {
            G.slot_of_pipe[k] = G.slot_of_pipe[k - 1];
        } else {
            G.slot_of_pipe[k] = (int)first_pipe.size();
            first_pipe.push_back(k);
        }
    }
    const int nu = (int)first_pipe.size();
    // group distinct slots by softening (first-appearance order of values);
    // each group is one force launch
    std::vector<double> values;
    std::vector<int> group_of(nu);
    std::unordered_map<uint64_t, int> group_of_bits;
    // kira passes one softening for every pipe: a run of equal values needs
    // no hash lookup (same groups, same order)
    uint64_t last_bits = 0;
    int last_group = -1;
    for (int u = 0; u < nu; ++u) {
        const double e = eps2[first_pipe[u]];
        uint64_t bits;
        memcpy(&bits, &e, sizeof bits);
        if (last_group >= 0 && bits == last_bits) {
            group_of[u] = last_group;
            continue;
        }
        auto it = group_of_bits.find(bits);
        if (it == group_of_bits.end()) {
            it = group_of_bits.emplace(bits, (int)values.size()).first;
            values.push_back(e);
        }
        group_of[u] = it->second;
        last_bits = bits;
        last_group = it->second;
    }
    std::vector<int> count(values.size(), 0), start(values.size(), 0);
    for (int u = 0; u < nu; ++u) ++count[group_of[u]];
    G.groups.clear();
    for (size_t g = 0, off = 0; g < values.size(); ++g) {
        start[g] = (int)off;
        G.groups.push_back({(int)off, count[g], 2.0 * values[g], true});
        off += count[g];
    }
    std::vector<int> pos(nu);
    for (int u = 0; u < nu; ++u) pos[u] = start[group_of[u]]++;
    G.xs.resize(3 * (size_t)nu);
    G.vs.resize(3 * (size_t)nu);
    G.h2s.resize(nu);
    G.e2s.resize(nu);
    for (int u = 0; u < nu; ++u) {
        const int k = first_pipe[u], p = pos[u];
        for (int c = 0; c < 3; ++c) {
            G.xs[3 * (size_t)p + c] = xi[k][c];
            G.vs[3 * (size_t)p + c] = vi[k][c];
        }
        G.h2s[p] = 2.0 * h2[k];
        G.e2s[p] = 2.0 * eps2[k];
    }
    for (int k = 0; k < ni; ++k) G.slot_of_pipe[k] = pos[G.slot_of_pipe[k]];
    G.n_pipes = ni;
    G.n_slots = nu;
    return calc_begin("h3calc_firsthalf", nj, nu,
                      reinterpret_cast<const double(*)[3]>(G.xs.data()),
                      reinterpret_cast<const double(*)[3]>(G.vs.data()), nullptr, G.h2s.data(),
                      G.e2s.data(), G.groups.data(), (int)G.groups.size(), 1);
}

int lasthalf(int ni, double (*acc)[3], double (*jerk)[3], double *pot)
{
    const Results *h = calc_wait("h3calc_lasthalf");
    if (!h) return -1;
    if (ni > G.n_pipes)
        return fail(-1, "h3calc_lasthalf: ni = %d but the first half had %d pipes", ni,
                    G.n_pipes);
    for (int k = 0; k < ni; ++k) {
        const int p = G.slot_of_pipe[k];
        for (int c = 0; c < 3; ++c) {
            acc[k][c] = h->acc[3 * (size_t)p + c];
            jerk[k][c] = h->jerk[3 * (size_t)p + c];
        }
        pot[k] = h->pot[p];
    }
    return 0;
}

// Build every list of the last i-block; a list longer than the store makes
// the store grow to the longest list and the lists are built again, so
// h3nblist_ always returns complete lists.
int nbread()
{
    int rc = read_nl();
    if (rc == 1) {
        long long want = 1;
        while (want < nl_max_count()) want <<= 1;
        E.opt.nb_cap = (int)std::min<long long>(want, 1LL << 30);
        rc = read_nl();
    }
    return rc == 0 ? 0 : fail(-1, "h3nbread: neighbour lists failed (%d)", rc);
}

int nblist(int chip, int *nblist)
{
    if (!E.nb_valid) return fail(-1, "h3nblist: lists not read (call h3nbread_ first)");
    if (chip < 0 || chip >= G.n_pipes)
        return fail(-1, "h3nblist: chip %d outside [0, %d)", chip, G.n_pipes);
    const int p = G.slot_of_pipe[chip];
    const int c = nl_count(p);
    if (c > E.opt.nb_cap) return fail(-1, "h3nblist: list of chip %d incomplete", chip);
    if (c > 0) memcpy(nblist, nl_list(p), sizeof(int) * (size_t)c);
    return c;
}

}  // namespace
}  // namespace g6

using namespace g6;

#pragma GCC visibility push(default)
extern "C" {

void h3open_()
{
    g4_read_env();
    if (do_open(E.opt.device)) fatal();
}

void h3close_() { (void)do_close(); }

int h3npipe_()
{
    g4_read_env();
    return E.opt.npipes;
}

void h3setnboards_(int *) {}

int h3getnboards_() { return 1; }

unsigned int h3wait_() { return 0; }

void h3setti_(double *ti) { E.ti = *ti; }

void h3setmode_(int *, int *) {}

void h3setled_(int *, int *) {}

int h3jpmax_()
{
    g4_read_env();
    return G.jpmax;
}

void h3jpdma_indirect_(int *nj, int *hostindex, double xj[][3], double vj[][3],
                       double aj[][3], double jj[][3], double *mj, double *tj, int *)
{
    dma_copy(*nj, hostindex, xj, vj, aj, jj, mj, tj, 0);
    dma_start(0);
}

void h3mjpdma_indirect_(int *nj, int *hostindex, double xj[][3], double vj[][3],
                        double aj[][3], double jj[][3], double *mj, double *tj, int *,
                        int *buff_id)
{
    const int b = *buff_id >= 0 ? *buff_id : G.buff_used++;
    dma_copy(*nj, hostindex, xj, vj, aj, jj, mj, tj, b);
}

void h3mjpdma_start_(int *buff_id) { dma_start(*buff_id); }

void h3mjpdma_flush_()
{
    for (int b = 0; b < G.buff_used; ++b) dma_start(b);
    G.buff_used = 0;
}

void h3calc_firsthalf_(int *nj, int *ni, double xi[][3], double vi[][3], double *eps2,
                       double *h2)
{
    if (firsthalf(*nj, *ni, xi, vi, eps2, h2)) fatal();
}

void h3calc_lasthalf_(int *ni, double acc[][3], double jerk[][3], double *pot)
{
    if (lasthalf(*ni, acc, jerk, pot)) fatal();
}

void h3calc_(int *nj, int *ni, double xi[][3], double vi[][3], double *eps2, double *h2,
             double acc[][3], double jerk[][3], double *pot)
{
    h3calc_firsthalf_(nj, ni, xi, vi, eps2, h2);
    h3calc_lasthalf_(ni, acc, jerk, pot);
}

void h3nbread_(int *)
{
    if (nbread()) fatal();
}

int h3nblist_(int *, int *chip, int *nblist_out)
{
    const int c = nblist(*chip, nblist_out);
    if (c < 0) {
        fatal();
        return 0;
    }
    return c;
}

int count_nblist_low(int, int chip)
{
    if (!E.nb_valid || chip < 0 || chip >= G.n_pipes) return 0;
    return nl_count(G.slot_of_pipe[chip]);
}

void h3setdebuglevel_(int *) {}

void h3checkdmabuffers_() {}

void set_time_check_mode(int) {}

// ---------------------------------------------------------------------------
// Extensions (include/g4x.h)
// ---------------------------------------------------------------------------
long long g4x_errors(void) { return G.errors; }

const char *g4x_last_error(void) { return E.last_error.c_str(); }

int g4x_set_option(const char *key, long long value)
{
    g4_read_env();
    const std::string k(key ? key : "");
    if (k == "npipes") E.opt.npipes = (int)std::max(1LL, value);
    else if (k == "jpmax") G.jpmax = (int)std::max(1LL, value);
    else if (k == "nb_cap") E.opt.nb_cap = (int)std::max(1LL, value);
    else if (k == "exit_on_error") G.exit_on_error = (int)value;
    else if (k == "dedupe") G.dedupe = (int)value;
    else if (k == "timing") E.opt.timing = (int)value;
    else if (k == "fused_small") E.opt.fused_small = value != 0 ? 1 : 0;
    else return -1;
    return 0;
}

long long g4x_get_option(const char *key)
{
    g4_read_env();
    const std::string k(key ? key : "");
    if (k == "npipes") return E.opt.npipes;
    if (k == "jpmax") return G.jpmax;
    if (k == "nb_cap") return E.opt.nb_cap;
    if (k == "exit_on_error") return G.exit_on_error;
    if (k == "dedupe") return G.dedupe;
    if (k == "timing") return E.opt.timing;
    if (k == "fused_small") return E.opt.fused_small;
    if (k == "last_slots") return G.n_slots;
    if (k == "last_groups") return (long long)G.groups.size();
    return -1;
}

}  // extern "C"
#pragma GCC visibility pop
