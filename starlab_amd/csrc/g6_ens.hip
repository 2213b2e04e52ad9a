// g6_ens.hip -- batched ensemble engine (include/g6ens.h): many independent
// N-body systems integrated together with kira's Hermite scheme, each system
// either on kira's individual power-of-two block steps (the default) or on
// one shared power-of-two step.
//
// Device state lives in one particle pool (capacity from the HBM budget);
// each resident system owns a contiguous range [off, off + n) and a header
// slot.  One engine step is four kernels, captured in a HIP graph:
//   k_ens_predict   thread per particle: kira's predictor to the system's
//                   next time (include/_dyn_.h:290-295), double-single copy
//                   for pairs; in block mode it also lists the particles of
//                   the block (t + dt = the block time, hdyn_schedule.C:641-745)
//   k_ens_force     one wave per (system, 64-i chunk, j-split), against the
//                   system's own j; in block mode the i are the block's
//                   particles.  Precision per system:
//                   64 (default): FP64 pair law written in the CPU oracle's
//                   exact operation order (hdyn_ev.C:2019-2036; IEEE division
//                   and square root, no contraction), so with one j-split a
//                   trajectory is bitwise the oracle's;
//                   32: the libg6 pair law (FP32 pairs on double-single
//                   positions, 256-j FP32 tiles folded into FP64)
//   k_ens_correct   thread per particle (of the block): sum j-splits in order,
//                   Hermite corrector (hdyn_ev.C:1235-1250, 1717-1718),
//                   Aarseth step (hdyn_ev.C:627-655).  Shared mode: atomic
//                   min of the step candidate into the system header.  Block
//                   mode: the particle's own next step by kira's power-of-two
//                   rule (hdyn_ev.C:866-887), atomic min of its next time.
//   k_ens_finalize  thread per slot: shared mode t += dt and the power-of-two
//                   rule on the system's step; block mode t = the block time,
//                   next block time = the smallest t + dt.  First steps as
//                   set_first_timestep (hdyn_ev.C:377-410).
// Atomic min of a positive double through its bit pattern is order-free, and
// every sum runs in a fixed order, so a system's trajectory does not depend
// on what else is resident.  In block mode the block's particles are listed
// in whatever order the waves append them; each particle's arithmetic does
// not depend on its place in the list.
//
// FP32 pairs carry ~1e-7 relative noise per force evaluation.  The Aarseth
// criterion differences successive accelerations (at3 = 2(a0 - a1) + ...,
// the third derivative times dt^3/6), so at small steps that noise looks
// like a large third derivative and the step collapses.  Precision 32
// therefore subtracts the noise floor (kNoiseFP32 |a| per evaluation) from
// |bt2| and |at3| before applying the criterion; precision 64 uses kira's
// criterion unchanged.
#include "g6_internal.h"
#include "../../include/g6ens.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace g6 {
namespace ens {

enum Phase : int { FREE = 0, INIT = 1, STEP = 2, FINAL = 3, DONE = 4, FAILED = 5 };

struct SysHdr {
    double t, dt, t_end, eta, eps2, dt_max;
    unsigned long long pmin;      // bits of the smallest candidate step (positive double)
    long long steps;              // engine steps (shared mode) or blocks (block mode)
    int off, n, phase, nsplit, jchunk, prec;   // prec: 32 or 64
    int lane;                                  // 0, 4 heavy; 1 light; 2, 3 workgroup kernels
    int mode;                                  // 0 shared step, 1 block steps
    // block mode: the next block time, the bits of the smallest t + dt among
    // the particles (atomic min, +inf: none left), the block's size, the
    // particle steps so far, and a flag raised when a particle's step underflows
    double tnext;
    unsigned long long tmin;
    long long psteps;
    int nact, fail;
};
static_assert(sizeof(SysHdr) == 128, "SysHdr layout");
constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;   // +inf as bits

// FP64 predicted state for the precision-64 pair law (one scalar-cache line).
struct JPred64 {
    double x, y, z, m, vx, vy, vz, pad;
};

constexpr double kNoiseFP32 = 4.0e-7;   // relative force noise of the FP32 pair law

struct Pool {
    int cap = 0;
    double *x = nullptr, *v = nullptr, *a = nullptr, *j = nullptr;   // [3][cap]
    double *xp = nullptr, *vp = nullptr;                             // [3][cap]
    double *m = nullptr, *pot = nullptr, *pot0 = nullptr;            // [cap]
    double *pt = nullptr, *pdt = nullptr;                            // [cap] block mode: t, dt
    int *act = nullptr;                                              // [cap] block mode: the block
    JPred *jp = nullptr;                                             // [cap]
    JPred64 *jp64 = nullptr;                                         // [cap]
    double *part = nullptr;                                          // [max_split][7][cap]
    int max_split = 0;
};

struct Chunk { int slot, i0; };
struct FItem { int slot, i0, split, pad; };

constexpr double kOneThird = 1.0 / 3.0;
constexpr double kOne12 = 1.0 / 12.0;

// ---------------------------------------------------------------------------
// Per-particle pieces shared by the step kernels and the fused kernel, so
// both paths perform the same operations in the same order.
// ---------------------------------------------------------------------------
// predict_loworder, include/_dyn_.h:290-295
__device__ __forceinline__ void ens_predict_one(const double x[3], const double v[3],
                                                const double a[3], const double j[3], double dt,
                                                double xp[3], double vp[3])
{
#pragma clang fp contract(off)
    const double dt3 = dt * kOneThird, dt2 = dt * 0.5;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xp[k] = ((j[k] * dt3 + a[k]) * dt2 + v[k]) * dt + x[k];
        vp[k] = (j[k] * dt2 + a[k]) * dt + v[k];
    }
}

__device__ __forceinline__ JPred make_jpred(const double xp[3], const double vp[3], double m)
{
#pragma clang fp contract(off)
    JPred q;
    q.xh = (float)xp[0]; q.xl = (float)(xp[0] - (double)q.xh);
    q.yh = (float)xp[1]; q.yl = (float)(xp[1] - (double)q.yh);
    q.zh = (float)xp[2]; q.zl = (float)(xp[2] - (double)q.zh);
    q.vx = (float)vp[0]; q.vy = (float)vp[1]; q.vz = (float)vp[2];
    q.m = (float)m;
#pragma unroll
    for (int k = 0; k < 6; ++k) q.pad[k] = 0.f;
    return q;
}

__device__ __forceinline__ JPred64 make_jpred64(const double xp[3], const double vp[3], double m)
{
    JPred64 d;
    d.x = xp[0]; d.y = xp[1]; d.z = xp[2]; d.m = m;
    d.vx = vp[0]; d.vy = vp[1]; d.vz = vp[2]; d.pad = 0.0;
    return d;
}

// Wave-wide minimum of a non-negative double (+inf where none), then one
// atomic min of its bit pattern into *dst (order-free for non-negative values).
__device__ __forceinline__ void wave_atomic_min(unsigned long long *dst, double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0 && v < __builtin_inf())
        atomicMin(dst, (unsigned long long)__double_as_longlong(v));
}

// Block mode, STEP: is particle p in the block at the system's next time?
__device__ __forceinline__ bool in_block(const Pool &P, const SysHdr &h, size_t p)
{
    const double ti = P.pt[p];
    return ti < h.t_end && ti + P.pdt[p] == h.tnext;
}

__global__ __launch_bounds__(64) void k_ens_predict(Pool P, SysHdr *H,
                                                    const Chunk *__restrict__ chunks,
                                                    const int *__restrict__ nchunks)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    SysHdr &h = H[ch.slot];
    if (h.phase < INIT || h.phase > FINAL) return;   // uniform in the wave
    const int i = ch.i0 + threadIdx.x;
    const bool valid = i < h.n;
    const size_t p = (size_t)h.off + (valid ? i : 0);
    const size_t cap = P.cap;
    double dt = h.phase == STEP ? h.dt : 0.0;
    if (h.mode == 1 && h.phase == STEP) {
        // every particle is predicted to the block time; the block's
        // particles are appended to the block list, the others' next times
        // go into the system's minimum
        const bool blk = valid && in_block(P, h, p);
        const double ti = P.pt[p];
        dt = h.tnext - ti;
        const unsigned long long bal = __ballot(blk);
        int base = 0;
        if (threadIdx.x == 0 && bal) base = atomicAdd(&h.nact, __popcll(bal));
        base = __shfl(base, 0);
        if (blk) P.act[(size_t)h.off + base + __popcll(bal & ((1ull << threadIdx.x) - 1))] = i;
        const double nxt = (valid && !blk && ti < h.t_end) ? ti + P.pdt[p] : __builtin_inf();
        wave_atomic_min(&h.tmin, nxt);
    }
    if (!valid) return;
    double x[3], v[3], a[3], jk[3], xp[3], vp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = P.x[k * cap + p]; v[k] = P.v[k * cap + p];
        a[k] = P.a[k * cap + p]; jk[k] = P.j[k * cap + p];
    }
    ens_predict_one(x, v, a, jk, dt, xp, vp);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        P.xp[k * cap + p] = xp[k];
        P.vp[k * cap + p] = vp[k];
    }
    P.jp[p] = make_jpred(xp, vp, P.m[p]);
    P.jp64[p] = make_jpred64(xp, vp, P.m[p]);
}

// ---------------------------------------------------------------------------
struct EAcc {
    float ax, ay, az, jx, jy, jz, kx, ky, kz, p;
};

template <bool MASK>
__device__ __forceinline__ void ens_pair(const JPred &J, int jj, unsigned rel, const JPred &I,
                                         float eps2, EAcc &a)
{
#pragma clang fp contract(off)
    const float dx = (J.xh - I.xh) + (J.xl - I.xl);
    const float dy = (J.yh - I.yh) + (J.yl - I.yl);
    const float dz = (J.zh - I.zh) + (J.zl - I.zl);
    // ((eps2 + dz*dz) + dy*dy) + dx*dx in three fused steps, as libg6's
    // force kernel evaluates it
    float r2 = __builtin_fmaf(dz, dz, eps2);
    r2 = __builtin_fmaf(dy, dy, r2);
    r2 = __builtin_fmaf(dx, dx, r2);
    const float dvx = J.vx - I.vx, dvy = J.vy - I.vy, dvz = J.vz - I.vz;
    float rv = dz * dvz;
    rv = __builtin_fmaf(dy, dvy, rv);
    rv = __builtin_fmaf(dx, dvx, rv);
    float rinv = __builtin_amdgcn_rsqf(r2);
    if (MASK) rinv = ((unsigned)jj == rel) ? 0.0f : rinv;
    const float rinv2 = rinv * rinv;
    const float mr = J.m * rinv;
    const float mr3 = mr * rinv2;
    const float q = mr3 * (rv * rinv2);
    a.p = a.p - mr;
    a.ax = __builtin_fmaf(mr3, dx, a.ax);
    a.ay = __builtin_fmaf(mr3, dy, a.ay);
    a.az = __builtin_fmaf(mr3, dz, a.az);
    a.jx = __builtin_fmaf(mr3, dvx, a.jx);
    a.jy = __builtin_fmaf(mr3, dvy, a.jy);
    a.jz = __builtin_fmaf(mr3, dvz, a.jz);
    a.kx = __builtin_fmaf(q, dx, a.kx);
    a.ky = __builtin_fmaf(q, dy, a.ky);
    a.kz = __builtin_fmaf(q, dz, a.kz);
}

template <bool MASK>
__device__ __forceinline__ void ens_tile(const JPred *__restrict__ jp, int tn, unsigned rel,
                                         const JPred &I, float eps2, EAcc &a)
{
    int jj = 0;
    for (; jj + 4 <= tn; jj += 4) {
        const JPred *g = jp + jj;
        ens_pair<MASK>(g[0], jj + 0, rel, I, eps2, a);
        ens_pair<MASK>(g[1], jj + 1, rel, I, eps2, a);
        ens_pair<MASK>(g[2], jj + 2, rel, I, eps2, a);
        ens_pair<MASK>(g[3], jj + 3, rel, I, eps2, a);
    }
    for (; jj < tn; ++jj) ens_pair<MASK>(jp[jj], jj, rel, I, eps2, a);
}

// FP64 pair law in the CPU oracle's operation order (hdyn_ev.C:2019-2036 as
// restated in oracle/g6_oracle.c), accumulating j in ascending order.
struct DAcc {
    double ax, ay, az, jx, jy, jz, p;
};

template <bool MASK>
__device__ __forceinline__ void ens_pair64(const JPred64 &J, int jj, int self, const JPred64 &I,
                                           double eps2, DAcc &a)
{
#pragma clang fp contract(off)
    if (MASK && jj == self) return;
    const double dx = J.x - I.x, dy = J.y - I.y, dz = J.z - I.z;
    const double du = J.vx - I.vx, dv = J.vy - I.vy, dw = J.vz - I.vz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double r2inv = 1.0 / (d2 + eps2);
    const double a3 = -3.0 * (dx * du + dy * dv + dz * dw) * r2inv;
    const double mrinv = J.m * sqrt(r2inv);
    a.p -= mrinv;
    const double mr3inv = mrinv * r2inv;
    a.ax += mr3inv * dx;
    a.ay += mr3inv * dy;
    a.az += mr3inv * dz;
    a.jx += mr3inv * (du + a3 * dx);
    a.jy += mr3inv * (dv + a3 * dy);
    a.jz += mr3inv * (dw + a3 * dz);
}

// The i of lane `lane` of work item chunk i0: position pos = i0 + lane of the
// system's i-list (its particles in order, or in block mode the block list).
// Lanes past the list's end get i = -1 and read particle 0.
struct ISel {
    int pos, i, ic;
};

__device__ __forceinline__ ISel select_i(const Pool &P, const SysHdr &h, bool blk, int cnt,
                                         int i0, int lane)
{
    ISel s;
    s.pos = i0 + lane;
    s.i = s.pos < cnt ? (blk ? P.act[(size_t)h.off + s.pos] : s.pos) : -1;
    s.ic = s.i >= 0 ? s.i : 0;
    return s;
}

// block mode, STEP: the i-list is the block; otherwise every particle
__device__ __forceinline__ bool block_step(const SysHdr &h) { return h.mode == 1 && h.phase == STEP; }

__device__ void ens_force64(const Pool &P, const SysHdr &h, const FItem &it, const ISel &sel)
{
    const int n = h.n;
    const int i = sel.i;
    const int ic = sel.ic;
    // the record address is the same in every lane: readfirstlane says so,
    // which keeps the j records on scalar loads into SGPRs
    const JPred64 *__restrict__ base = P.jp64 + __builtin_amdgcn_readfirstlane(h.off);
    const JPred64 I = base[ic];
    const int self = i;
    const int jbeg = __builtin_amdgcn_readfirstlane(it.split * h.jchunk);
    const int jend = __builtin_amdgcn_readfirstlane(min(jbeg + h.jchunk, n));
    const double eps2 = h.eps2;
    DAcc a;
    a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.p = 0.0;
    for (int t = jbeg; t < jend; t += kTile) {
        const int tn = min(kTile, jend - t);
        const unsigned rel = (unsigned)(self - t);
        if (__ballot(rel < (unsigned)tn) == 0) {
            for (int jj = 0; jj < tn; ++jj) ens_pair64<false>(base[t + jj], t + jj, self, I, eps2, a);
        } else {
            for (int jj = 0; jj < tn; ++jj) ens_pair64<true>(base[t + jj], t + jj, self, I, eps2, a);
        }
    }
    if (i >= 0) {
        const size_t cap = P.cap;
        const size_t p = (size_t)h.off + sel.pos;   // partials by list position
        double *out = P.part + (size_t)it.split * 7 * cap;
        out[0 * cap + p] = a.ax;
        out[1 * cap + p] = a.ay;
        out[2 * cap + p] = a.az;
        out[3 * cap + p] = a.jx;
        out[4 * cap + p] = a.jy;
        out[5 * cap + p] = a.jz;
        out[6 * cap + p] = a.p;
    }
}

__global__ __launch_bounds__(64) void k_ens_force(Pool P, const SysHdr *__restrict__ H,
                                                  const FItem *__restrict__ items,
                                                  const int *__restrict__ nitems)
{
    const int w = blockIdx.x;
    if (w >= *nitems) return;
    const FItem it = items[w];
    const SysHdr &h = H[it.slot];
    if (h.phase < INIT || h.phase > FINAL) return;
    const bool blk = block_step(h);
    const int cnt = blk ? h.nact : h.n;
    if (it.i0 >= cnt) return;                      // uniform in the wave
    const int lane = threadIdx.x;
    const ISel sel = select_i(P, h, blk, cnt, it.i0, lane);
    if (h.prec == 64) {
        ens_force64(P, h, it, sel);
        return;
    }
    const int n = h.n;
    const int i = sel.i;
    const int ic = sel.ic;
    // uniform record address (readfirstlane): the j records stay on scalar loads
    const JPred *__restrict__ base = P.jp + __builtin_amdgcn_readfirstlane(h.off);
    const JPred I = base[ic];
    const int self = i;
    const float eps2 = (float)h.eps2;
    const int jbeg = __builtin_amdgcn_readfirstlane(it.split * h.jchunk);
    const int jend = __builtin_amdgcn_readfirstlane(min(jbeg + h.jchunk, n));
    double A[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int t = jbeg; t < jend; t += kTile) {
        const int tn = min(kTile, jend - t);
        EAcc a;
        a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.kx = a.ky = a.kz = a.p = 0.f;
        const unsigned rel = (unsigned)(self - t);
        if (__ballot(rel < (unsigned)tn) == 0)
            ens_tile<false>(base + t, tn, rel, I, eps2, a);
        else
            ens_tile<true>(base + t, tn, rel, I, eps2, a);
        A[0] += (double)a.ax;
        A[1] += (double)a.ay;
        A[2] += (double)a.az;
        A[3] += (double)a.jx - 3.0 * (double)a.kx;
        A[4] += (double)a.jy - 3.0 * (double)a.ky;
        A[5] += (double)a.jz - 3.0 * (double)a.kz;
        A[6] += (double)a.p;
    }
    if (i >= 0) {
        const size_t cap = P.cap;
        const size_t p = (size_t)h.off + sel.pos;   // partials by list position
        double *out = P.part + (size_t)it.split * 7 * cap;
#pragma unroll
        for (int c = 0; c < 7; ++c) out[c * cap + p] = A[c];
    }
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(const double *u, const double *v)
{
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

// One particle's state as the corrector sees it.
struct PState {
    double x[3], v[3], a[3], j[3], pot, pot0;
};

// Corrector and step candidate for one particle, given its predicted state
// and the new force; updates the fields the phase writes and returns the
// candidate step (+inf: none).
__device__ __forceinline__ double ens_correct_one(PState &st, const double xp[3],
                                                  const double vp[3], const double a1[3],
                                                  const double j1[3], double pot, int phase,
                                                  double dt, double eta, int prec)
{
#pragma clang fp contract(off)
    double cand = __builtin_inf();
    if (phase == INIT) {
        // set_first_timestep, hdyn_ev.C:377-410: eta/16 * sqrt(min(a^2/j^2, -pot/a^2))
        for (int k = 0; k < 3; ++k) { st.a[k] = a1[k]; st.j[k] = j1[k]; }
        st.pot = pot;
        st.pot0 = pot;
        const double eta_init = eta / 16;
        const double asq = dot3(a1, a1), jsq = dot3(j1, j1);
        double dt_adot = 1e300, dt_ff = 1e300;
        if (jsq > 0) dt_adot = asq / jsq;
        if (pot < 0 && asq > 0) dt_ff = -pot / asq;
        cand = eta_init * sqrt(fmin(dt_adot, dt_ff));
        cand = fmin(cand, eta_init);
    } else if (phase == STEP) {
        double bt2[3], at3[3];
        for (int k = 0; k < 3; ++k) {
            const double a0 = st.a[k], j0 = st.j[k];
            // get_derivs, hdyn_ev.C:1235-1250
            bt2[k] = -3 * (a0 - a1[k]) - dt * (2 * j0 + j1[k]);
            at3[k] = 2 * (a0 - a1[k]) + dt * (j0 + j1[k]);
        }
        for (int k = 0; k < 3; ++k) {
            // hdyn_ev.C:1717-1718
            st.x[k] = xp[k] + (0.05 * at3[k] + kOne12 * bt2[k]) * dt * dt;
            st.v[k] = vp[k] + (0.25 * at3[k] + kOneThird * bt2[k]) * dt;
            st.a[k] = a1[k];
            st.j[k] = j1[k];
        }
        st.pot = pot;
        // Aarseth criterion, hdyn_ev.C:627-655
        const double dt2 = dt * dt, dt4 = dt2 * dt2, dt6 = dt4 * dt2;
        double l2 = 36 * dot3(at3, at3) / dt6;
        double k2 = 4 * dot3(bt2, bt2) / dt4;
        if (prec != 64) {
            // remove the FP32 force noise: a0 - a1 carries ~2 kNoiseFP32 |a|
            const double na = kNoiseFP32 * sqrt(dot3(a1, a1));
            const double at3m = fmax(sqrt(dot3(at3, at3)) - 4 * na, 0.0);
            const double bt2m = fmax(sqrt(dot3(bt2, bt2)) - 6 * na, 0.0);
            l2 = 36 * at3m * at3m / dt6;
            k2 = 4 * bt2m * bt2m / dt4;
        }
        const double j2 = dot3(j1, j1);
        const double asq = dot3(a1, a1);
        cand = eta * sqrt((sqrt(asq * k2) + j2) / (sqrt(j2 * l2) + k2));
        if (!(cand > 0)) cand = __builtin_inf();
    } else {   // FINAL: potential at the exact final positions
        st.pot = pot;
    }
    return cand;
}

__device__ __forceinline__ double pow2_below(double x)
{
    int e;
    const double f = frexp(x, &e);   // x = f * 2^e, f in [0.5, 1)
    (void)f;
    return ldexp(1.0, e - 1);
}

// Block mode: the particle's next step after a corrected step of length T
// ending at block time tb, given the criterion's value nat: halved once if
// nat falls short, doubled when tb is commensurate with 2T and 2T stays within
// dt_max (hdyn_ev.C:866-887), cut to end at t_end.
__device__ __forceinline__ double block_next_step(double nat, double T, double tb, double dt_max,
                                                  double t_end)
{
#pragma clang fp contract(off)
    double nd = T;
    if (nat < T) nd = 0.5 * T;
    else if (nat >= 2 * T && fmod(tb, 2 * T) == 0.0 && 2 * T <= dt_max) nd = 2 * T;
    if (tb + nd > t_end) nd = t_end - tb;
    return nd;
}

// Block mode: a particle's time and step after its correction (INIT: the
// first step from the set_first_timestep candidate; STEP: the next step);
// returns its next time, +inf once it has reached t_end.  *fail is set when
// a step underflows dt_min.
__device__ __forceinline__ double block_advance(double &ti, double &di, double cand, int phase,
                                                const SysHdr &h, double dt_min, int *fail)
{
    if (phase == INIT) {
        ti = 0.0;
        di = pow2_below(fmin(cand, h.dt_max));
        if (!(di > dt_min)) *fail = 1;
    } else if (phase == STEP) {
        const double tb = h.tnext;
        di = block_next_step(cand, di, tb, h.dt_max, h.t_end);
        ti = tb;
        if (tb < h.t_end && !(di > dt_min)) *fail = 1;
    }
    return ti < h.t_end ? ti + di : __builtin_inf();
}

// The split partials of list position q, added in split order: split 0 as
// is (no "0 +" rounding), then splits 1, 2, ...  The loads of four splits
// are issued before their additions, so a system of many splits waits for
// memory once per four splits instead of once per split.
__device__ __forceinline__ void sum_splits(const Pool &P, int nsplit, size_t q, double a1[3],
                                           double j1[3], double &pot)
{
#pragma clang fp contract(off)
    constexpr int U = 4;
    const size_t cap = P.cap;
    double A[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) A[c] = P.part[c * cap + q];
    for (int s0 = 1; s0 < nsplit; s0 += U) {
        double v[U][7];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double *in = P.part + (size_t)min(s0 + u, nsplit - 1) * 7 * cap;
#pragma unroll
            for (int c = 0; c < 7; ++c) v[u][c] = in[c * cap + q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s0 + u >= nsplit) break;
#pragma unroll
            for (int c = 0; c < 7; ++c) A[c] += v[u][c];
        }
    }
    a1[0] = A[0]; a1[1] = A[1]; a1[2] = A[2];
    j1[0] = A[3]; j1[1] = A[4]; j1[2] = A[5];
    pot = A[6];
}

__global__ __launch_bounds__(64) void k_ens_correct(Pool P, SysHdr *__restrict__ H,
                                                    const Chunk *__restrict__ chunks,
                                                    const int *__restrict__ nchunks,
                                                    double dt_min)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    SysHdr &h = H[ch.slot];
    const int phase = h.phase;
    if (phase < INIT || phase > FINAL) return;
    const bool blk = block_step(h);
    const int cnt = blk ? h.nact : h.n;
    if (ch.i0 >= cnt) return;                      // uniform in the wave
    const ISel sel = select_i(P, h, blk, cnt, ch.i0, threadIdx.x);
    const int i = sel.i;
    double cand = __builtin_inf(), next = __builtin_inf();
    int fail = 0;
    if (i >= 0) {
        const size_t cap = P.cap;
        const size_t p = (size_t)h.off + i;
        const size_t q = (size_t)h.off + sel.pos;   // partials by list position
        double a1[3], j1[3], pot;
        sum_splits(P, h.nsplit, q, a1, j1, pot);
        PState st;
        double xp[3], vp[3];
        for (int k = 0; k < 3; ++k) {
            st.x[k] = P.x[k * cap + p]; st.v[k] = P.v[k * cap + p];
            st.a[k] = P.a[k * cap + p]; st.j[k] = P.j[k * cap + p];
            xp[k] = P.xp[k * cap + p]; vp[k] = P.vp[k * cap + p];
        }
        st.pot = 0; st.pot0 = 0;
        double ti = 0, di = 0;
        if (h.mode == 1) { ti = P.pt[p]; di = P.pdt[p]; }
        cand = ens_correct_one(st, xp, vp, a1, j1, pot, phase, h.mode == 1 ? di : h.dt, h.eta,
                               h.prec);
        if (h.mode == 1 && phase != FINAL) {
            next = block_advance(ti, di, cand, phase, h, dt_min, &fail);
            P.pt[p] = ti;
            P.pdt[p] = di;
        }
        if (phase == INIT) {
            for (int k = 0; k < 3; ++k) { P.a[k * cap + p] = st.a[k]; P.j[k * cap + p] = st.j[k]; }
            P.pot[p] = st.pot;
            P.pot0[p] = st.pot0;
        } else if (phase == STEP) {
            for (int k = 0; k < 3; ++k) {
                P.x[k * cap + p] = st.x[k];
                P.v[k * cap + p] = st.v[k];
                P.a[k * cap + p] = st.a[k];
                P.j[k * cap + p] = st.j[k];
            }
            P.pot[p] = st.pot;
        } else {
            P.pot[p] = st.pot;
        }
    }
    if (h.mode == 1) {
        // the block's next times join the system's minimum; a failed step
        // marks the system
        wave_atomic_min(&h.tmin, next);
        if (__ballot(fail) && threadIdx.x == 0) atomicOr(&h.fail, 1);
        return;
    }
    // wave minimum, then one atomic per wave
    if (phase != FINAL) wave_atomic_min(&h.pmin, cand);
}

// t += dt and kira's step rules for one system, given the smallest step
// candidate `nat` of its particles.
__device__ __forceinline__ void ens_finalize_hdr(SysHdr &h, double nat, long long max_steps,
                                                 double dt_min)
{
    if (h.phase == INIT) {
        double dt = fmin(nat, h.dt_max);
        dt = pow2_below(dt);
        while (h.t != 0 && fmod(h.t, dt) != 0) dt *= 0.5;
        h.dt = dt;
        h.phase = (dt > dt_min) ? STEP : FAILED;
        if (h.t + h.dt > h.t_end) h.dt = h.t_end - h.t;
    } else if (h.phase == STEP) {
        h.t += h.dt;
        h.steps += 1;
        if (h.t >= h.t_end) {
            h.phase = FINAL;
        } else {
            // hdyn_ev.C:866-887: halve at most once; double only when the
            // time is commensurate with the doubled step
            const double T = h.dt;
            double nd = T;
            if (nat < T) nd = 0.5 * T;
            else if (nat >= 2 * T && fmod(h.t, 2 * T) == 0.0 && 2 * T <= h.dt_max) nd = 2 * T;
            if (h.t + nd > h.t_end) nd = h.t_end - h.t;
            h.dt = nd;
            if (!(nd > dt_min) || h.steps > max_steps) h.phase = FAILED;
        }
    } else if (h.phase == FINAL) {
        h.phase = DONE;
    }
}

// Block mode: the system moves to the block just done (STEP) and on to the
// next block time tmin, the smallest t + dt of its particles (+inf once every
// particle is at t_end: then the final potential pass, FINAL, and DONE).
// nact = the block's size, fail = a particle's step underflowed.
__device__ __forceinline__ void ens_finalize_block(SysHdr &h, double tmin, int nact, int fail,
                                                   long long max_steps)
{
    if (h.phase == INIT) {
        h.t = 0.0;
        h.tnext = tmin;
        h.dt = tmin;
        h.phase = (fail || !(tmin < __builtin_inf())) ? FAILED : STEP;
    } else if (h.phase == STEP) {
        h.dt = h.tnext - h.t;
        h.t = h.tnext;
        h.steps += 1;
        h.psteps += nact;
        if (fail || h.steps > max_steps) {
            h.phase = FAILED;
        } else if (!(tmin < __builtin_inf())) {
            h.phase = FINAL;
        } else {
            h.tnext = tmin;
        }
    } else if (h.phase == FINAL) {
        h.phase = DONE;
    }
}

__global__ void k_ens_finalize(SysHdr *__restrict__ H, int nslots, int *__restrict__ phase_out,
                               long long max_steps, double dt_min, int lane,
                               int *__restrict__ bcount)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) bcount[0] = bcount[1] = 0;   // the lane's block lists, for the next step
    if (s >= nslots) return;
    SysHdr &h = H[s];
    if (h.lane != lane || h.phase < INIT || h.phase > FINAL) return;
    if (h.mode == 1) {
        ens_finalize_block(h, __longlong_as_double((long long)h.tmin), h.nact, h.fail, max_steps);
        h.tmin = kInfBits;
        h.nact = 0;
    } else {
        const double nat = __longlong_as_double((long long)h.pmin);
        ens_finalize_hdr(h, nat, max_steps, dt_min);
    }
    h.pmin = kInfBits;
    phase_out[s] = h.phase;
}

// ---------------------------------------------------------------------------
// Block-step path with prediction inside the force kernel (option
// block_fusion = 1, the default).  Per engine step:
//   k_ens_bselect  thread per particle: reads only t and dt, lists the block
//                  (t + dt = the block time) and takes the others' next times
//                  into the system's minimum;
//   k_ens_bitems   thread per block-mode system: appends the system's work to
//                  the lane's device lists, (system, 64-i chunk) for the
//                  corrector and (system, chunk, j-split) for the force, with
//                  one atomic per wave;
//   k_ens_bforce   grid-stride over the force items: each wave predicts its
//                  j-split 64 j at a time (one per lane) into LDS and sums the
//                  pairs from there.  Precision 64, and precision 32 for
//                  blocks of more than 32 i: one lane per i, in the same order
//                  as k_ens_force on predicted records, so results are bitwise
//                  equal.  Precision 32 blocks of at most 32 i (most blocks
//                  hold a few particles): NIP i-slots (the next power of two)
//                  of Q = 64 / NIP lanes each; lane q of a slot takes the j
//                  congruent to q mod Q, folds its FP32 sums of each 256-j
//                  tile into FP64, and the slot's lanes add their sums in a
//                  fixed butterfly.  The order depends only on the block's
//                  size, so a system's trajectory still depends only on its
//                  own particles;
//   k_ens_bcorrect grid-stride over the chunks: sums the splits, recomputes
//                  the particle's own prediction, corrects it and takes its
//                  next step;
//   k_ens_finalize as above; it also clears the lane's list counters.
// No prediction is written to HBM, and waves never start on empty work.
// ---------------------------------------------------------------------------
struct BLists {
    Chunk *chunks;
    FItem *items;
    int *count;        // [0] chunks, [1] items of this step
};

__global__ __launch_bounds__(64) void k_ens_bselect(Pool P, SysHdr *H,
                                                    const Chunk *__restrict__ chunks,
                                                    const int *__restrict__ nchunks)
{
    const int c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    SysHdr &h = H[ch.slot];
    if (h.mode != 1 || h.phase != STEP) return;   // INIT / FINAL: every particle, in order
    const int lane = threadIdx.x;
    const int i = ch.i0 + lane;
    const bool valid = i < h.n;
    const size_t p = (size_t)h.off + (valid ? i : 0);
    const double ti = P.pt[p];
    const bool blk = valid && ti < h.t_end && ti + P.pdt[p] == h.tnext;
    const unsigned long long bal = __ballot(blk);
    int base = 0;
    if (lane == 0 && bal) base = atomicAdd(&h.nact, __popcll(bal));
    base = __shfl(base, 0);
    if (blk) P.act[(size_t)h.off + base + __popcll(bal & ((1ull << lane) - 1))] = i;
    const double nxt = (valid && !blk && ti < h.t_end) ? ti + P.pdt[p] : __builtin_inf();
    wave_atomic_min(&h.tmin, nxt);
}

// Inclusive prefix sum over the wave.
__device__ __forceinline__ int wave_scan(int v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// Thread per block-mode system (64-thread workgroups): the system's work
// items, placed by a prefix sum over the wave and one atomic per wave.
__global__ __launch_bounds__(64) void k_ens_bitems(const SysHdr *__restrict__ H,
                                                   const int *__restrict__ sys,
                                                   const int *__restrict__ nsys, BLists L)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    const int lane = threadIdx.x;
    int slot = -1, nc = 0, ns = 0;
    if (s < *nsys) {
        slot = sys[s];
        const SysHdr &h = H[slot];
        if (h.mode == 1 && h.phase >= INIT && h.phase <= FINAL) {
            const int cnt = h.phase == STEP ? h.nact : h.n;
            nc = (cnt + 63) / 64;
            ns = h.nsplit;
        }
    }
    const int pc = wave_scan(nc), pi = wave_scan(nc * ns);
    int bc = 0, bi = 0;
    if (lane == 63) {
        if (pc) bc = atomicAdd(&L.count[0], pc);
        if (pi) bi = atomicAdd(&L.count[1], pi);
    }
    bc = __shfl(bc, 63);
    bi = __shfl(bi, 63);
    const int c0 = bc + pc - nc, w0 = bi + pi - nc * ns;
    for (int k = 0; k < nc; ++k) {
        L.chunks[c0 + k] = Chunk{slot, 64 * k};
        for (int sp = 0; sp < ns; ++sp) L.items[w0 + k * ns + sp] = FItem{slot, 64 * k, sp, 0};
    }
}

// Particle p's state predicted to the system's current evaluation time
// (the block time in STEP, its own time otherwise).
__device__ __forceinline__ void bpredict(const Pool &P, const SysHdr &h, size_t p, double xp[3],
                                         double vp[3])
{
#pragma clang fp contract(off)
    const size_t cap = P.cap;
    double x[3], v[3], a[3], jk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = P.x[k * cap + p]; v[k] = P.v[k * cap + p];
        a[k] = P.a[k * cap + p]; jk[k] = P.j[k * cap + p];
    }
    const double dt = h.phase == STEP ? h.tnext - P.pt[p] : 0.0;
    ens_predict_one(x, v, a, jk, dt, xp, vp);
}

__global__ __launch_bounds__(64) void k_ens_bforce(Pool P, const SysHdr *__restrict__ H,
                                                   const FItem *__restrict__ items,
                                                   const int *__restrict__ nitems)
{
#pragma clang fp contract(off)
    // 64 predicted j: JPred64 records (FP64), JPred records (FP32, one lane
    // per i: every lane reads the same record), or the ten JPred fields as
    // separate arrays (FP32 packed: lanes read different j, which then sit
    // in different banks)
    __shared__ double rec[64 * 8];
    JPred64 *r64 = reinterpret_cast<JPred64 *>(rec);
    JPred *r32 = reinterpret_cast<JPred *>(rec);
    float (*f32)[64] = reinterpret_cast<float (*)[64]>(rec);
    const int lane = threadIdx.x;
    const int total = *nitems;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const FItem it = items[w];
        const SysHdr &h = H[it.slot];
        const bool blk = h.phase == STEP;
        const int cnt = blk ? h.nact : h.n;
        // i-slots: one per lane, or for FP32 blocks of at most 32 i NIP slots
        // of Q lanes.  The choice depends on the block's size, not on the
        // chunk: the block list is filled in no fixed order, so a particle's
        // chunk is not reproducible, while its arithmetic must be.
        int nip = 64;
        if (h.prec != 64 && cnt <= 32) {
            nip = 1;
            while (nip < cnt) nip *= 2;
        }
        const int Q = 64 / nip;
        const int slot = lane / Q, q = lane - slot * Q;
        const ISel sel = select_i(P, h, blk, cnt, it.i0, slot);
        const int self = sel.i;
        const size_t off = (size_t)h.off;
        double xi[3], vi[3];
        bpredict(P, h, off + sel.ic, xi, vi);
        const double mi = P.m[off + sel.ic];
        const int jbeg = it.split * h.jchunk;
        const int jend = min(jbeg + h.jchunk, h.n);
        double A[7] = {0, 0, 0, 0, 0, 0, 0};
        if (h.prec == 64) {
            const JPred64 I = make_jpred64(xi, vi, mi);
            DAcc a;
            a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.p = 0.0;
            for (int t = jbeg; t < jend; t += 64) {
                const int tn = min(64, jend - t);
                __syncthreads();                   // the previous records have been read
                if (lane < tn) {
                    double xp[3], vp[3];
                    bpredict(P, h, off + t + lane, xp, vp);
                    r64[lane] = make_jpred64(xp, vp, P.m[off + t + lane]);
                }
                __syncthreads();
                const unsigned rel = (unsigned)(self - t);
                if (__ballot(rel < (unsigned)tn) == 0) {
                    for (int jj = 0; jj < tn; ++jj) ens_pair64<false>(r64[jj], t + jj, self, I, h.eps2, a);
                } else {
                    for (int jj = 0; jj < tn; ++jj) ens_pair64<true>(r64[jj], t + jj, self, I, h.eps2, a);
                }
            }
            A[0] = a.ax; A[1] = a.ay; A[2] = a.az;
            A[3] = a.jx; A[4] = a.jy; A[5] = a.jz; A[6] = a.p;
        } else {
            const JPred I = make_jpred(xi, vi, mi);
            const float eps2 = (float)h.eps2;
            for (int t = jbeg; t < jend; t += kTile) {
                const int tn = min(kTile, jend - t);
                EAcc e;
                e.ax = e.ay = e.az = e.jx = e.jy = e.jz = e.kx = e.ky = e.kz = e.p = 0.f;
                for (int s0 = 0; s0 < tn; s0 += 64) {
                    const int sn = min(64, tn - s0);
                    __syncthreads();
                    if (lane < sn) {
                        double xp[3], vp[3];
                        const size_t qj = off + t + s0 + lane;
                        bpredict(P, h, qj, xp, vp);
                        const JPred r = make_jpred(xp, vp, P.m[qj]);
                        if (Q == 1) {
                            r32[lane] = r;
                        } else {
                            f32[0][lane] = r.xh; f32[1][lane] = r.yh; f32[2][lane] = r.zh;
                            f32[3][lane] = r.xl; f32[4][lane] = r.yl; f32[5][lane] = r.zl;
                            f32[6][lane] = r.vx; f32[7][lane] = r.vy; f32[8][lane] = r.vz;
                            f32[9][lane] = r.m;
                        }
                    }
                    __syncthreads();
                    // one lane per i: the tile's pairs in ascending order, as
                    // ens_tile; Q lanes per i: lane q takes the j congruent to
                    // q mod Q
                    const unsigned rel = (unsigned)(self - (t + s0));
                    const bool mask = __ballot(rel < (unsigned)sn) != 0;
                    if (Q == 1) {
                        if (!mask) ens_tile<false>(r32, sn, rel, I, eps2, e);
                        else ens_tile<true>(r32, sn, rel, I, eps2, e);
                    } else {
                        auto rec_at = [&](int jj) {
                            JPred r;
                            r.xh = f32[0][jj]; r.yh = f32[1][jj]; r.zh = f32[2][jj];
                            r.xl = f32[3][jj]; r.yl = f32[4][jj]; r.zl = f32[5][jj];
                            r.vx = f32[6][jj]; r.vy = f32[7][jj]; r.vz = f32[8][jj];
                            r.m = f32[9][jj];
                            return r;
                        };
                        if (!mask) {
                            for (int jj = q; jj < sn; jj += Q) ens_pair<false>(rec_at(jj), jj, rel, I, eps2, e);
                        } else {
                            for (int jj = q; jj < sn; jj += Q) ens_pair<true>(rec_at(jj), jj, rel, I, eps2, e);
                        }
                    }
                }
                A[0] += (double)e.ax;
                A[1] += (double)e.ay;
                A[2] += (double)e.az;
                A[3] += (double)e.jx - 3.0 * (double)e.kx;
                A[4] += (double)e.jy - 3.0 * (double)e.ky;
                A[5] += (double)e.jz - 3.0 * (double)e.kz;
                A[6] += (double)e.p;
            }
            // butterfly over a slot's Q lanes (a + b and b + a round alike, so
            // every lane of the slot ends with the same sums)
            for (int o = 1; o < Q; o <<= 1) {
#pragma unroll
                for (int c = 0; c < 7; ++c) A[c] += __shfl_xor(A[c], o);
            }
        }
        if (q == 0 && self >= 0) {
            const size_t cap = P.cap;
            const size_t qo = off + sel.pos;   // partials by list position
            double *out = P.part + (size_t)it.split * 7 * cap;
#pragma unroll
            for (int c = 0; c < 7; ++c) out[c * cap + qo] = A[c];
        }
    }
}

// Block path: correct the particle at position pos of the system's i-list
// (particle i): sum its split partials in order, recompute its prediction,
// correct it, give it its next step.  Returns its next time (+inf: at t_end
// or FINAL); *fail is set when its step underflows.
__device__ __forceinline__ double bcorrect_one(const Pool &P, const SysHdr &h, int i, int pos,
                                               double dt_min, int *fail)
{
#pragma clang fp contract(off)
    const int phase = h.phase;
    const size_t cap = P.cap;
    const size_t p = (size_t)h.off + i;
    const size_t q = (size_t)h.off + pos;
    double a1[3], j1[3], pot;
    sum_splits(P, h.nsplit, q, a1, j1, pot);
    PState st;
    double xp[3], vp[3];
    bpredict(P, h, p, xp, vp);
    for (int k = 0; k < 3; ++k) {
        st.x[k] = P.x[k * cap + p]; st.v[k] = P.v[k * cap + p];
        st.a[k] = P.a[k * cap + p]; st.j[k] = P.j[k * cap + p];
    }
    st.pot = 0; st.pot0 = 0;
    double ti = P.pt[p], di = P.pdt[p];
    const double cand = ens_correct_one(st, xp, vp, a1, j1, pot, phase, di, h.eta, h.prec);
    double next = __builtin_inf();
    if (phase != FINAL) {
        next = block_advance(ti, di, cand, phase, h, dt_min, fail);
        P.pt[p] = ti;
        P.pdt[p] = di;
    }
    if (phase == INIT) {
        for (int k = 0; k < 3; ++k) { P.a[k * cap + p] = st.a[k]; P.j[k * cap + p] = st.j[k]; }
        P.pot[p] = st.pot;
        P.pot0[p] = st.pot0;
    } else if (phase == STEP) {
        for (int k = 0; k < 3; ++k) {
            P.x[k * cap + p] = st.x[k];
            P.v[k * cap + p] = st.v[k];
            P.a[k * cap + p] = st.a[k];
            P.j[k * cap + p] = st.j[k];
        }
        P.pot[p] = st.pot;
    } else {
        P.pot[p] = st.pot;
    }
    return next;
}

__global__ __launch_bounds__(64) void k_ens_bcorrect(Pool P, SysHdr *H,
                                                     const Chunk *__restrict__ chunks,
                                                     const int *__restrict__ nchunks,
                                                     double dt_min)
{
    const int total = *nchunks;
    for (int c = blockIdx.x; c < total; c += gridDim.x) {
        const Chunk ch = chunks[c];
        SysHdr &h = H[ch.slot];
        const bool blk = h.phase == STEP;
        const int cnt = blk ? h.nact : h.n;
        const ISel sel = select_i(P, h, blk, cnt, ch.i0, threadIdx.x);
        double next = __builtin_inf();
        int fail = 0;
        if (sel.i >= 0) next = bcorrect_one(P, h, sel.i, sel.pos, dt_min, &fail);
        wave_atomic_min(&h.tmin, next);
        if (__ballot(fail) && threadIdx.x == 0) atomicOr(&h.fail, 1);
    }
}

// ---------------------------------------------------------------------------
// Fused kernel for light systems (n <= 64): one wave owns one system, one
// lane per particle, and runs whole engine steps (predict, force, correct,
// step rules) in a loop of up to max_iters steps, leaving early when the
// system is done.  Predicted j live in LDS; state stays in registers and is
// written back at the end.  The arithmetic is the step kernels' (shared
// helpers above, one j-split, j ascending), so a system's trajectory is the
// same bit for bit on either path.  In block mode every lane predicts its
// particle to the block time and sums its force; only the block's lanes
// correct and take new steps.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_ens_fused(Pool P, SysHdr *__restrict__ H,
                                                  const Chunk *__restrict__ list,
                                                  const int *__restrict__ nlist,
                                                  int *__restrict__ phase_out,
                                                  long long max_steps, double dt_min,
                                                  int max_iters)
{
#pragma clang fp contract(off)
    __shared__ JPred64 s64[64];
    __shared__ JPred s32[64];
    const int w = blockIdx.x;
    if (w >= *nlist) return;
    const int slot = list[w].slot;
    SysHdr h = H[slot];
    if (h.phase < INIT || h.phase > FINAL || h.n > 64) return;
    const int n = h.n;
    const int i = threadIdx.x;
    const bool act = i < n;
    const int ic = act ? i : n - 1;
    const size_t cap = P.cap;
    const size_t p = (size_t)h.off + ic;
    const double m = P.m[p];
    PState st;
    for (int k = 0; k < 3; ++k) {
        st.x[k] = P.x[k * cap + p]; st.v[k] = P.v[k * cap + p];
        st.a[k] = P.a[k * cap + p]; st.j[k] = P.j[k * cap + p];
    }
    st.pot = P.pot[p];
    st.pot0 = P.pot0[p];
    const int self = act ? i : -1;
    const bool bm = h.mode == 1;
    double ti = 0, di = 0;   // block mode: the particle's time and step
    if (bm) { ti = P.pt[p]; di = P.pdt[p]; }
    double xp[3], vp[3];
    for (int it = 0; it < max_iters; ++it) {
        const int phase = h.phase;
        double dpred = phase == STEP ? h.dt : 0.0;
        bool inb = act;
        if (bm && phase == STEP) {
            inb = act && ti < h.t_end && ti + di == h.tnext;
            dpred = h.tnext - ti;
        }
        ens_predict_one(st.x, st.v, st.a, st.j, dpred, xp, vp);
        if (act) {
            s64[i] = make_jpred64(xp, vp, m);
            s32[i] = make_jpred(xp, vp, m);
        }
        __syncthreads();
        double a1[3], j1[3], pot;
        if (h.prec == 64) {
            const JPred64 I = s64[ic];
            DAcc acc;
            acc.ax = acc.ay = acc.az = acc.jx = acc.jy = acc.jz = acc.p = 0.0;
            for (int jj = 0; jj < n; ++jj) ens_pair64<true>(s64[jj], jj, self, I, h.eps2, acc);
            a1[0] = acc.ax; a1[1] = acc.ay; a1[2] = acc.az;
            j1[0] = acc.jx; j1[1] = acc.jy; j1[2] = acc.jz;
            pot = acc.p;
        } else {
            const JPred I = s32[ic];
            EAcc e;
            e.ax = e.ay = e.az = e.jx = e.jy = e.jz = e.kx = e.ky = e.kz = e.p = 0.f;
            ens_tile<true>(s32, n, (unsigned)self, I, (float)h.eps2, e);
            double A[7] = {0, 0, 0, 0, 0, 0, 0};
            A[0] += (double)e.ax;
            A[1] += (double)e.ay;
            A[2] += (double)e.az;
            A[3] += (double)e.jx - 3.0 * (double)e.kx;
            A[4] += (double)e.jy - 3.0 * (double)e.ky;
            A[5] += (double)e.jz - 3.0 * (double)e.kz;
            A[6] += (double)e.p;
            a1[0] = A[0]; a1[1] = A[1]; a1[2] = A[2];
            j1[0] = A[3]; j1[1] = A[4]; j1[2] = A[5];
            pot = A[6];
        }
        __syncthreads();   // LDS is rewritten by the next step's prediction
        if (bm) {
            double next = __builtin_inf();
            int fail = 0;
            if (inb) {
                const double c = ens_correct_one(st, xp, vp, a1, j1, pot, phase, di, h.eta, h.prec);
                if (phase != FINAL) next = block_advance(ti, di, c, phase, h, dt_min, &fail);
            } else if (act && ti < h.t_end) {
                next = ti + di;
            }
            for (int off = 32; off > 0; off >>= 1) next = fmin(next, __shfl_xor(next, off));
            const int nact = phase == STEP ? __popcll(__ballot(inb)) : 0;
            const int anyfail = __ballot(fail) != 0;
            ens_finalize_block(h, next, nact, anyfail, max_steps);   // every lane, same values
            if (h.phase == DONE || h.phase == FAILED) break;
            continue;
        }
        double cand = __builtin_inf();
        if (act) cand = ens_correct_one(st, xp, vp, a1, j1, pot, phase, h.dt, h.eta, h.prec);
        for (int off = 32; off > 0; off >>= 1) cand = fmin(cand, __shfl_xor(cand, off));
        ens_finalize_hdr(h, cand, max_steps, dt_min);   // every lane, same values
        if (h.phase == DONE || h.phase == FAILED) break;
    }
    if (act) {
        for (int k = 0; k < 3; ++k) {
            P.x[k * cap + p] = st.x[k]; P.v[k * cap + p] = st.v[k];
            P.a[k * cap + p] = st.a[k]; P.j[k * cap + p] = st.j[k];
            P.xp[k * cap + p] = xp[k]; P.vp[k * cap + p] = vp[k];
        }
        P.pot[p] = st.pot;
        P.pot0[p] = st.pot0;
        if (bm) { P.pt[p] = ti; P.pdt[p] = di; }
    }
    if (i == 0) {
        h.pmin = kInfBits;
        h.tmin = kInfBits;
        h.nact = 0;
        H[slot] = h;
        phase_out[slot] = h.phase;
    }
}

// ---------------------------------------------------------------------------
// Workgroup-fused kernel for medium systems (light_max < n <= wg_max_n): one
// workgroup of NT threads owns one system and runs whole engine steps in a
// loop of up to max_iters steps, with workgroup barriers between the phases
// instead of kernel boundaries:
//   predict  thread t predicts particles t, t + NT, ... into a 64-B record
//            per particle in LDS (JPred for precision 32, JPred64 for 64);
//   force    the same thread sums each of its i over the system's j-splits
//            and 256-j tiles in the step kernels' order (split totals added
//            in split order, as k_ens_correct does);
//   correct  the step kernels' corrector, then a workgroup minimum of the
//            step candidates and the header update by thread 0.
// Each step reproduces the step kernels' operations, so a trajectory is the
// same bit for bit as on the graph path.  Dynamic LDS: n_max records of 64 B,
// then NT / 64 doubles for the minimum.  Systems whose records do not fit
// stay on the graph path.
// ---------------------------------------------------------------------------
static_assert(sizeof(JPred) == 64 && sizeof(JPred64) == 64, "LDS record size");
constexpr int kWgLargeThreads = 1024;   // threads per system of the large lane

size_t ens_wg_lds(int nmax, int nt) { return (size_t)nmax * 64 + (size_t)(nt / 64) * 8; }

template <int NT, int PREC>
__global__ __launch_bounds__(NT) void k_ens_wg(Pool P, SysHdr *__restrict__ H,
                                               const Chunk *__restrict__ list, int nlist,
                                               int *__restrict__ phase_out, long long max_steps,
                                               double dt_min, int max_iters, int nmax)
{
#pragma clang fp contract(off)
    extern __shared__ double wg_lds[];
    __shared__ SysHdr sh;
    const int w = blockIdx.x;
    if (w >= nlist) return;                        // uniform per workgroup
    const int slot = list[w].slot;
    const int tid = threadIdx.x;
    if (tid == 0) sh = H[slot];
    __syncthreads();
    if (sh.phase < INIT || sh.phase > FINAL || sh.n > nmax || sh.prec != PREC) return;   // uniform
    const int n = sh.n;
    const size_t cap = P.cap;
    const size_t off = (size_t)sh.off;
    JPred *rec32 = reinterpret_cast<JPred *>(wg_lds);
    JPred64 *rec64 = reinterpret_cast<JPred64 *>(wg_lds);
    double *red = wg_lds + (size_t)nmax * 8;
    const int rounds = (n + NT - 1) / NT;
    for (int it = 0; it < max_iters; ++it) {
        const int phase = sh.phase;
        const double dt = phase == STEP ? sh.dt : 0.0;
        constexpr int prec = PREC;
        // ---- predict ----
        for (int i = tid; i < n; i += NT) {
            const size_t p = off + i;
            double x[3], v[3], a[3], jk[3], xp[3], vp[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = P.x[k * cap + p]; v[k] = P.v[k * cap + p];
                a[k] = P.a[k * cap + p]; jk[k] = P.j[k * cap + p];
            }
            ens_predict_one(x, v, a, jk, dt, xp, vp);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                P.xp[k * cap + p] = xp[k];
                P.vp[k * cap + p] = vp[k];
            }
            const double m = P.m[p];
            if (prec == 64) rec64[i] = make_jpred64(xp, vp, m);
            else rec32[i] = make_jpred(xp, vp, m);
        }
        __syncthreads();
        // ---- force + correct, NT i at a time ----
        double cand = __builtin_inf();
        for (int r = 0; r < rounds; ++r) {
            const int i = r * NT + tid;
            const bool act = i < n;
            const int ic = act ? i : n - 1;
            const int self = act ? i : -1;
            double a1[3], j1[3], pot = 0.0;
            for (int sp = 0; sp < sh.nsplit; ++sp) {
                const int jbeg = sp * sh.jchunk;
                const int jend = min(jbeg + sh.jchunk, n);
                double A[7] = {0, 0, 0, 0, 0, 0, 0};
                if (prec == 64) {
                    const JPred64 I = rec64[ic];
                    DAcc d;
                    d.ax = d.ay = d.az = d.jx = d.jy = d.jz = d.p = 0.0;
                    for (int t = jbeg; t < jend; t += kTile) {
                        const int tn = min(kTile, jend - t);
                        const unsigned rel = (unsigned)(self - t);
                        if (__ballot(rel < (unsigned)tn) == 0) {
                            for (int jj = 0; jj < tn; ++jj)
                                ens_pair64<false>(rec64[t + jj], t + jj, self, I, sh.eps2, d);
                        } else {
                            for (int jj = 0; jj < tn; ++jj)
                                ens_pair64<true>(rec64[t + jj], t + jj, self, I, sh.eps2, d);
                        }
                    }
                    A[0] = d.ax; A[1] = d.ay; A[2] = d.az;
                    A[3] = d.jx; A[4] = d.jy; A[5] = d.jz; A[6] = d.p;
                } else {
                    const JPred I = rec32[ic];
                    const float eps2 = (float)sh.eps2;
                    for (int t = jbeg; t < jend; t += kTile) {
                        const int tn = min(kTile, jend - t);
                        EAcc e;
                        e.ax = e.ay = e.az = e.jx = e.jy = e.jz = e.kx = e.ky = e.kz = e.p = 0.f;
                        const unsigned rel = (unsigned)(self - t);
                        if (__ballot(rel < (unsigned)tn) == 0)
                            ens_tile<false>(rec32 + t, tn, rel, I, eps2, e);
                        else
                            ens_tile<true>(rec32 + t, tn, rel, I, eps2, e);
                        A[0] += (double)e.ax;
                        A[1] += (double)e.ay;
                        A[2] += (double)e.az;
                        A[3] += (double)e.jx - 3.0 * (double)e.kx;
                        A[4] += (double)e.jy - 3.0 * (double)e.ky;
                        A[5] += (double)e.jz - 3.0 * (double)e.kz;
                        A[6] += (double)e.p;
                    }
                }
                // split 0 as is, later splits added in order (k_ens_correct)
                if (sp == 0) {
                    a1[0] = A[0]; a1[1] = A[1]; a1[2] = A[2];
                    j1[0] = A[3]; j1[1] = A[4]; j1[2] = A[5];
                    pot = A[6];
                } else {
                    a1[0] += A[0]; a1[1] += A[1]; a1[2] += A[2];
                    j1[0] += A[3]; j1[1] += A[4]; j1[2] += A[5];
                    pot += A[6];
                }
            }
            if (act) {
                const size_t p = off + i;
                PState st;
                double xp[3], vp[3];
                for (int k = 0; k < 3; ++k) {
                    st.x[k] = P.x[k * cap + p]; st.v[k] = P.v[k * cap + p];
                    st.a[k] = P.a[k * cap + p]; st.j[k] = P.j[k * cap + p];
                    xp[k] = P.xp[k * cap + p]; vp[k] = P.vp[k * cap + p];
                }
                st.pot = 0; st.pot0 = 0;
                const double c = ens_correct_one(st, xp, vp, a1, j1, pot, phase, dt, sh.eta, prec);
                cand = fmin(cand, c);
                if (phase == INIT) {
                    for (int k = 0; k < 3; ++k) { P.a[k * cap + p] = st.a[k]; P.j[k * cap + p] = st.j[k]; }
                    P.pot[p] = st.pot;
                    P.pot0[p] = st.pot0;
                } else if (phase == STEP) {
                    for (int k = 0; k < 3; ++k) {
                        P.x[k * cap + p] = st.x[k];
                        P.v[k * cap + p] = st.v[k];
                        P.a[k * cap + p] = st.a[k];
                        P.j[k * cap + p] = st.j[k];
                    }
                    P.pot[p] = st.pot;
                } else {
                    P.pot[p] = st.pot;
                }
            }
        }
        // ---- workgroup minimum of the step candidates, header update ----
        for (int o = 32; o > 0; o >>= 1) cand = fmin(cand, __shfl_xor(cand, o));
        if ((tid & 63) == 0) red[tid >> 6] = cand;
        __syncthreads();
        if (tid == 0) {
            double nat = red[0];
            for (int q = 1; q < NT / 64; ++q) nat = fmin(nat, red[q]);
            ens_finalize_hdr(sh, nat, max_steps, dt_min);
        }
        __syncthreads();    // header visible; LDS records free for the next step
        if (sh.phase == DONE || sh.phase == FAILED) break;
    }
    if (tid == 0) {
        sh.pmin = kInfBits;
        H[slot] = sh;
        phase_out[slot] = sh.phase;
    }
}

// ---------------------------------------------------------------------------
// Workgroup kernel for block-step systems (option wgb_steps > 0): one
// workgroup of NT threads (256 up to wg_small_max particles, else 512) owns
// one system of light_max < n <= wg_max_n and
// runs up to max_iters block steps per launch:
//   select   the block (t + dt = the block time) into an LDS list, and the
//            smallest next time of the other particles;
//   predict  every j to the block time into LDS (FP64 records, or the ten
//            FP32 fields as separate arrays);
//   force    the waves take the fused block path's work items (64-i chunk,
//            j-split) in turn and compute each as k_ens_bforce does (same
//            lane packing, pair order, tile folds and butterfly), writing the
//            split partials to HBM;
//   correct  bcorrect_one per block particle, a workgroup minimum of the next
//            times, and the header update by thread 0.
// Each step performs the fused block path's operations on the same values,
// so a trajectory is bitwise that of the step graphs.  The step costs no
// kernel launch and no pass over other systems: a block of a few particles
// costs a prediction of the system's j and a few thousand pairs.  Dynamic
// LDS: the records of nmax j, then nmax ints for the block list.
// ---------------------------------------------------------------------------
// threads per system of the large block-step lane: 512 leave 256 VGPRs per
// thread (1024 threads allow 128, and the kernel spilled there)
constexpr int kWgbLargeThreads = 512;

size_t ens_wgb_lds(int nmax, int prec)
{
    const size_t rec = prec == 64 ? sizeof(JPred64) : 10 * sizeof(float);
    return (size_t)nmax * (rec + sizeof(int));
}

template <int NT, int PREC>
__global__ __launch_bounds__(NT) void k_ens_wgb(Pool P, SysHdr *__restrict__ H,
                                                const Chunk *__restrict__ list, int nlist,
                                                int *__restrict__ phase_out, long long max_steps,
                                                double dt_min, int max_iters, int nmax)
{
#pragma clang fp contract(off)
    constexpr int NW = NT / 64;
    extern __shared__ double wgb_lds[];
    __shared__ SysHdr sh;
    __shared__ int s_cnt, s_fail;
    __shared__ double s_min[NW];
    const int w = blockIdx.x;
    if (w >= nlist) return;                        // uniform per workgroup
    const int slot = list[w].slot;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) sh = H[slot];
    __syncthreads();
    if (sh.mode != 1 || sh.phase < INIT || sh.phase > FINAL || sh.n > nmax || sh.prec != PREC)
        return;                                    // uniform
    const int n = sh.n;
    const size_t off = (size_t)sh.off;
    JPred64 *r64 = reinterpret_cast<JPred64 *>(wgb_lds);
    float *f32 = reinterpret_cast<float *>(wgb_lds);   // [10][nmax]
    int *blist = reinterpret_cast<int *>(reinterpret_cast<char *>(wgb_lds) +
                                         (size_t)nmax * (PREC == 64 ? sizeof(JPred64) : 40));
    auto rec32 = [&](int j) {
        JPred r;
        r.xh = f32[0 * nmax + j]; r.yh = f32[1 * nmax + j]; r.zh = f32[2 * nmax + j];
        r.xl = f32[3 * nmax + j]; r.yl = f32[4 * nmax + j]; r.zl = f32[5 * nmax + j];
        r.vx = f32[6 * nmax + j]; r.vy = f32[7 * nmax + j]; r.vz = f32[8 * nmax + j];
        r.m = f32[9 * nmax + j];
        return r;
    };
    for (int it = 0; it < max_iters; ++it) {
        const int phase = sh.phase;
        const bool blk = phase == STEP;
        if (tid == 0) { s_cnt = 0; s_fail = 0; }
        __syncthreads();
        // ---- the block, and the others' next times ----
        double nmin = __builtin_inf();
        if (blk) {
            for (int i0 = 0; i0 < n; i0 += NT) {
                const int i = i0 + tid;
                const bool valid = i < n;
                const size_t p = off + (valid ? i : 0);
                const double ti = P.pt[p], di = P.pdt[p];
                const bool in = valid && ti < sh.t_end && ti + di == sh.tnext;
                const unsigned long long bal = __ballot(in);
                int base = 0;
                if (lane == 0 && bal) base = atomicAdd(&s_cnt, __popcll(bal));
                base = __shfl(base, 0);
                if (in) blist[base + __popcll(bal & ((1ull << lane) - 1))] = i;
                if (valid && !in && ti < sh.t_end) nmin = fmin(nmin, ti + di);
            }
        }
        // ---- every j predicted to the evaluation time ----
        for (int j = tid; j < n; j += NT) {
            double xp[3], vp[3];
            bpredict(P, sh, off + j, xp, vp);
            if (PREC == 64) {
                r64[j] = make_jpred64(xp, vp, P.m[off + j]);
            } else {
                const JPred r = make_jpred(xp, vp, P.m[off + j]);
                f32[0 * nmax + j] = r.xh; f32[1 * nmax + j] = r.yh; f32[2 * nmax + j] = r.zh;
                f32[3 * nmax + j] = r.xl; f32[4 * nmax + j] = r.yl; f32[5 * nmax + j] = r.zl;
                f32[6 * nmax + j] = r.vx; f32[7 * nmax + j] = r.vy; f32[8 * nmax + j] = r.vz;
                f32[9 * nmax + j] = r.m;
            }
        }
        __syncthreads();
        const int cnt = blk ? s_cnt : n;
        // ---- force: the fused block path's items, one wave each ----
        {
            const int nc = (cnt + 63) / 64, ns = sh.nsplit;
            int nip = 64;
            if (PREC != 64 && cnt <= 32) {
                nip = 1;
                while (nip < cnt) nip *= 2;
            }
            const int Q = 64 / nip;
            const int islot = lane / Q, q = lane - islot * Q;
            for (int item = wave; item < nc * ns; item += NW) {
                const int c = item / ns, sp = item - c * ns;
                const int pos = 64 * c + islot;
                const int self = pos < cnt ? (blk ? blist[pos] : pos) : -1;
                const int ic = self >= 0 ? self : 0;
                const int jbeg = sp * sh.jchunk;
                const int jend = min(jbeg + sh.jchunk, n);
                double A[7] = {0, 0, 0, 0, 0, 0, 0};
                if (PREC == 64) {
                    const JPred64 I = r64[ic];
                    DAcc a;
                    a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.p = 0.0;
                    for (int t = jbeg; t < jend; t += 64) {
                        const int tn = min(64, jend - t);
                        const unsigned rel = (unsigned)(self - t);
                        if (__ballot(rel < (unsigned)tn) == 0) {
                            for (int jj = 0; jj < tn; ++jj)
                                ens_pair64<false>(r64[t + jj], t + jj, self, I, sh.eps2, a);
                        } else {
                            for (int jj = 0; jj < tn; ++jj)
                                ens_pair64<true>(r64[t + jj], t + jj, self, I, sh.eps2, a);
                        }
                    }
                    A[0] = a.ax; A[1] = a.ay; A[2] = a.az;
                    A[3] = a.jx; A[4] = a.jy; A[5] = a.jz; A[6] = a.p;
                } else {
                    const JPred I = rec32(ic);
                    const float eps2 = (float)sh.eps2;
                    for (int t = jbeg; t < jend; t += kTile) {
                        const int tn = min(kTile, jend - t);
                        EAcc e;
                        e.ax = e.ay = e.az = e.jx = e.jy = e.jz = e.kx = e.ky = e.kz = e.p = 0.f;
                        for (int s0 = 0; s0 < tn; s0 += 64) {
                            const int sn = min(64, tn - s0);
                            const int b0 = t + s0;
                            const unsigned rel = (unsigned)(self - b0);
                            if (__ballot(rel < (unsigned)sn) == 0) {
                                for (int jj = q; jj < sn; jj += Q)
                                    ens_pair<false>(rec32(b0 + jj), jj, rel, I, eps2, e);
                            } else {
                                for (int jj = q; jj < sn; jj += Q)
                                    ens_pair<true>(rec32(b0 + jj), jj, rel, I, eps2, e);
                            }
                        }
                        A[0] += (double)e.ax;
                        A[1] += (double)e.ay;
                        A[2] += (double)e.az;
                        A[3] += (double)e.jx - 3.0 * (double)e.kx;
                        A[4] += (double)e.jy - 3.0 * (double)e.ky;
                        A[5] += (double)e.jz - 3.0 * (double)e.kz;
                        A[6] += (double)e.p;
                    }
                    for (int o = 1; o < Q; o <<= 1) {
#pragma unroll
                        for (int k = 0; k < 7; ++k) A[k] += __shfl_xor(A[k], o);
                    }
                }
                if (q == 0 && self >= 0) {
                    const size_t cap = P.cap;
                    double *out = P.part + (size_t)sp * 7 * cap;
#pragma unroll
                    for (int k = 0; k < 7; ++k) out[k * cap + off + pos] = A[k];
                }
            }
        }
        __syncthreads();                           // the partials are written
        // ---- correct the block, next block time ----
        int fail = 0;
        for (int pos = tid; pos < cnt; pos += NT) {
            const int i = blk ? blist[pos] : pos;
            nmin = fmin(nmin, bcorrect_one(P, sh, i, pos, dt_min, &fail));
        }
        for (int o = 32; o > 0; o >>= 1) nmin = fmin(nmin, __shfl_xor(nmin, o));
        if (lane == 0) s_min[wave] = nmin;
        if (__ballot(fail) && lane == 0) atomicOr(&s_fail, 1);
        __syncthreads();
        if (tid == 0) {
            double m = s_min[0];
            for (int k = 1; k < NW; ++k) m = fmin(m, s_min[k]);
            ens_finalize_block(sh, m, blk ? cnt : 0, s_fail, max_steps);
        }
        __syncthreads();                           // header visible; LDS free for the next step
        if (sh.phase == DONE || sh.phase == FAILED) break;
    }
    if (tid == 0) {
        sh.pmin = sh.tmin = kInfBits;
        sh.nact = 0;
        sh.fail = 0;
        H[slot] = sh;
        phase_out[slot] = sh.phase;
    }
}

// ---------------------------------------------------------------------------
// admission and retirement
// staged particle record: m, x[3], v[3], dest index (as double)
__global__ void k_ens_unpack(Pool P, const double *__restrict__ rec, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *r = rec + 8 * (size_t)k;
    const size_t p = (size_t)r[7];
    const size_t cap = P.cap;
    P.m[p] = r[0];
    for (int c = 0; c < 3; ++c) {
        P.x[c * cap + p] = r[1 + c];
        P.v[c * cap + p] = r[4 + c];
        P.a[c * cap + p] = 0.0;
        P.j[c * cap + p] = 0.0;
    }
    P.pot[p] = 0.0;
    P.pot0[p] = 0.0;
    P.pt[p] = 0.0;
    P.pdt[p] = 0.0;
}

__global__ void k_ens_set_headers(SysHdr *__restrict__ H, const SysHdr *__restrict__ staged,
                                  const int *__restrict__ slots, int n, int *__restrict__ phase_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    H[slots[k]] = staged[k];
    phase_out[slots[k]] = staged[k].phase;
}

// output record per particle: x[3], v[3], pot, pot0 (8 doubles)
__global__ void k_ens_pack(Pool P, const int *__restrict__ src, double *__restrict__ out, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t p = src[k];
    const size_t cap = P.cap;
    double *o = out + 8 * (size_t)k;
    for (int c = 0; c < 3; ++c) {
        o[c] = P.x[c * cap + p];
        o[3 + c] = P.v[c * cap + p];
    }
    o[6] = P.pot[p];
    o[7] = P.pot0[p];
}

// headers of the listed slots, gathered for one download
__global__ void k_ens_gather_headers(const SysHdr *__restrict__ H, const int *__restrict__ slots,
                                     SysHdr *__restrict__ out, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = H[slots[k]];
}

// ===========================================================================
// host engine
// ===========================================================================
namespace {

struct Job {
    long long id;
    int n;
    std::vector<double> m, x, v;
    double t_end, eta, eps2;
    int prec, mode;
};

struct SlotInfo {
    bool busy = false;
    long long job = -1;
    int off = 0, n = 0, nsplit = 1, lane = 0;
    std::vector<double> m, x0, v0;
    double eps2 = 0, t_end = 0;
    int prec = 64, mode = 1;
};

struct Result {
    int n;
    std::vector<double> x, v;
    double t, e0, e1;
    long long steps;
    bool failed;
};

struct GraphExec {
    hipGraphExec_t exec = nullptr;
    int chunk_grid = 0, item_grid = 0, steps = 0;
    int sel_grid = 0, sys_grid = 0;   // block path: select chunks, systems
};

// Systems run in lanes with their own work lists and streams:
//   lane 1  light systems (n <= light_max, e.g. three-body scatterings): one
//           wave per system in k_ens_fused;
//   lane 2  light_max < n <= wg_small_max: one 256-thread workgroup per
//           system in k_ens_wg;
//   lane 3  wg_small_max < n <= wg_max_n: one kWgLargeThreads workgroup per
//           system in k_ens_wg;
//   lane 0  the rest (and every non-light system when wg_steps = 0): the
//           four step kernels over (system, i-chunk, j-split) work items,
//           replayed as HIP graphs.
// The light lane keeps stepping while the other lanes run, so a light
// system's latency does not grow with the heavy step time.
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Chunk *d_chunks = nullptr, *h_chunks = nullptr;
    FItem *d_items = nullptr, *h_items = nullptr;
    int *d_counts = nullptr, *h_counts = nullptr;   // [0] chunks, [1] items
    int chunk_cap = 0, item_cap = 0;
    int n_chunks = 0, n_items = 0;
    std::vector<GraphExec> graphs;
    long long steps = 0;
    int nmax = 0;                 // largest system of a workgroup-fused lane
    // its systems by group, listed in this order: shared steps FP32, shared
    // FP64, block steps FP32, block FP64 (one launch per group)
    int ngrp[4] = {0, 0, 0, 0};
    // block-step systems on the fused block path: every 64-particle chunk (for
    // the select kernel); the step's chunk and item lists are built on the
    // device (d_bchunks, d_bitems, counters d_bcount)
    Chunk *d_schunks = nullptr, *h_schunks = nullptr;
    int *d_bsys = nullptr, *h_bsys = nullptr;
    Chunk *d_bchunks = nullptr;
    FItem *d_bitems = nullptr;
    int *d_bcount = nullptr;
    int *d_nsel = nullptr, *h_nsel = nullptr;   // [0] select chunks, [1] systems
    int bcap_chunks = 0, bcap_items = 0, bcap_sys = 0;
    int n_schunks = 0, n_bsys = 0, n_bitems_max = 0;
};

// kernel_ms slots, named kernel_us_<name> by g6e_get_option
enum KernelSlot { KS_PREDICT, KS_FORCE, KS_CORRECT, KS_BSELECT, KS_BITEMS, KS_BFORCE, KS_BCORRECT,
                  KS_FINALIZE, KS_FUSED, KS_N };
const char *const kKernelNames[KS_N] = {"predict", "force", "correct", "bselect", "bitems",
                                        "bforce", "bcorrect", "finalize", "fused"};

struct Engine {
    bool open = false;
    hipStream_t stream = nullptr;
    std::string err;
    Pool P;
    int max_slots = 0;
    SysHdr *d_hdr = nullptr;
    int *d_phase = nullptr, *h_phase = nullptr;
    // 0: heavy systems (step graphs); 1: light (n <= light_max); 2, 3:
    // workgroup-kernel systems; 4: a second heavy lane (option heavy_lanes)
    Lane lanes[5];
    bool items_dirty = true;
    // staging
    double *h_stage = nullptr, *d_stage = nullptr;
    size_t stage_cap = 0;             // doubles
    SysHdr *h_hstage = nullptr, *d_hstage = nullptr;
    int *h_hslots = nullptr, *d_hslots = nullptr;
    int *h_src = nullptr, *d_src = nullptr;
    double *h_out = nullptr, *d_out = nullptr;
    size_t out_cap = 0;               // particles
    std::vector<SlotInfo> slots;
    std::map<int, int> free_ranges;   // off -> len
    std::deque<Job> queue;
    long long next_id = 0;
    std::unordered_map<long long, Result> results;
    std::vector<long long> finished;
    // options
    int steps_per_graph = 8;
    int jsplit_target = 256;
    int max_split = 16;
    int lookahead = 64;
    int precision = 64;
    int block_steps = 1;              // 1: kira's individual block steps, 0: one shared step
    int block_fusion = 1;             // block steps: 1 = prediction inside the force kernel,
                                      // device-built work lists; 0 = the four step kernels
    int light_max = 64;
    int fused_steps = 256;            // light lane: steps per fused launch (0: graphs)
    int wg_steps = 0;                 // medium lanes: steps per workgroup-fused launch (0: graphs;
                                      // measured slower than the graphs on the 4096-cluster batch)
    int wg_small_max = 512;           // lane 2: light_max < n <= this, 256 threads per system
    int wg_max_n = 2048;              // lane 3: wg_small_max < n <= this, 1024 threads per system
    int heavy_lanes = 1;              // heavy systems split over lanes 0 and 4 when 2
    int wgb_steps = 0;                // block-step systems in the workgroup lanes: block steps
                                      // per launch of k_ens_wgb (0: the fused block path)
    double dt_max = 0.125;
    long long max_steps = 50000000;
    // stats
    long long engine_steps = 0, n_finished = 0;
    double pairs = 0, dev_seconds = 0;
    // host time of g6e_run by phase (seconds): retire, admit, rebuild, graphs,
    // waiting for the lanes, the phase download
    double host_s[6] = {0, 0, 0, 0, 0, 0};
    // option kernel_timing: lanes 0 and 1 launch their step kernels one by one
    // (no graphs) between HIP events; device milliseconds per kernel
    int kernel_timing = 0;
    double kernel_ms[KS_N] = {};
    hipEvent_t kev[2] = {nullptr, nullptr};
};

Engine X;

int efail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    X.err = buf;
    fprintf(stderr, "libg6 ensemble: %s\n", buf);
    return code;
}

#define ECHK(call)                                                             \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return efail(-1, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

template <class T> void dfree(T *&p) { if (p) { (void)hipFree(p); p = nullptr; } }
template <class T> void hfree(T *&p) { if (p) { (void)hipHostFree(p); p = nullptr; } }

size_t bytes_per_particle(int max_split)
{
    return sizeof(double) * (3 * 6 + 3 + 2) + sizeof(int) + sizeof(JPred) + sizeof(JPred64) +
           sizeof(double) * 7 * (size_t)max_split;
}

// Medium systems go to the workgroup lanes when their step mode's kernel is
// on: k_ens_wg (shared steps, wg_steps > 0) or k_ens_wgb (block steps,
// wgb_steps > 0).
int lane_for(int n, int mode)
{
    if (n <= X.light_max) return 1;
    const bool wg = mode == 0 ? X.wg_steps > 0 : X.wgb_steps > 0;
    if (wg && n <= X.wg_small_max) return 2;
    if (wg && n <= X.wg_max_n) return 3;
    return 0;
}

int nsplit_for(int n)
{
    if (X.jsplit_target <= 0) return 1;
    int s = (n + X.jsplit_target - 1) / X.jsplit_target;
    return std::max(1, std::min(s, X.max_split));
}

int ensure_stage(size_t doubles)
{
    if (doubles <= X.stage_cap) return 0;
    size_t cap = std::max(doubles, X.stage_cap * 2);
    hfree(X.h_stage); dfree(X.d_stage);
    X.stage_cap = 0;
    ECHK(hipHostMalloc((void **)&X.h_stage, sizeof(double) * cap, hipHostMallocDefault));
    ECHK(hipMalloc((void **)&X.d_stage, sizeof(double) * cap));
    X.stage_cap = cap;
    return 0;
}

int ensure_out(size_t particles)
{
    if (particles <= X.out_cap) return 0;
    size_t cap = std::max(particles, X.out_cap * 2);
    hfree(X.h_out); dfree(X.d_out); hfree(X.h_src); dfree(X.d_src);
    X.out_cap = 0;
    ECHK(hipHostMalloc((void **)&X.h_out, sizeof(double) * 8 * cap, hipHostMallocDefault));
    ECHK(hipMalloc((void **)&X.d_out, sizeof(double) * 8 * cap));
    ECHK(hipHostMalloc((void **)&X.h_src, sizeof(int) * cap, hipHostMallocDefault));
    ECHK(hipMalloc((void **)&X.d_src, sizeof(int) * cap));
    X.out_cap = cap;
    return 0;
}

void destroy_graphs(Lane &L)
{
    for (auto &g : L.graphs) if (g.exec) (void)hipGraphExecDestroy(g.exec);
    L.graphs.clear();
}

int ensure_items(Lane &L, int chunks, int items)
{
    if (chunks > L.chunk_cap) {
        int cap = std::max(chunks, L.chunk_cap * 2);
        hfree(L.h_chunks); dfree(L.d_chunks);
        ECHK(hipHostMalloc((void **)&L.h_chunks, sizeof(Chunk) * cap, hipHostMallocDefault));
        ECHK(hipMalloc((void **)&L.d_chunks, sizeof(Chunk) * cap));
        L.chunk_cap = cap;
        destroy_graphs(L);
    }
    if (items > L.item_cap) {
        int cap = std::max(items, L.item_cap * 2);
        hfree(L.h_items); dfree(L.d_items);
        ECHK(hipHostMalloc((void **)&L.h_items, sizeof(FItem) * cap, hipHostMallocDefault));
        ECHK(hipMalloc((void **)&L.d_items, sizeof(FItem) * cap));
        L.item_cap = cap;
        destroy_graphs(L);
    }
    return 0;
}

// Capacity of the lane's block-path lists: select chunks (= the most chunks a
// step can list), systems, and force items.
int ensure_blists(Lane &L, int chunks, int nsys, int items)
{
    if (chunks > L.bcap_chunks) {
        int cap = std::max(chunks, L.bcap_chunks * 2);
        hfree(L.h_schunks); dfree(L.d_schunks); dfree(L.d_bchunks);
        ECHK(hipHostMalloc((void **)&L.h_schunks, sizeof(Chunk) * cap, hipHostMallocDefault));
        ECHK(hipMalloc((void **)&L.d_schunks, sizeof(Chunk) * cap));
        ECHK(hipMalloc((void **)&L.d_bchunks, sizeof(Chunk) * cap));
        L.bcap_chunks = cap;
        destroy_graphs(L);
    }
    if (nsys > L.bcap_sys) {
        int cap = std::max(nsys, L.bcap_sys * 2);
        hfree(L.h_bsys); dfree(L.d_bsys);
        ECHK(hipHostMalloc((void **)&L.h_bsys, sizeof(int) * cap, hipHostMallocDefault));
        ECHK(hipMalloc((void **)&L.d_bsys, sizeof(int) * cap));
        L.bcap_sys = cap;
        destroy_graphs(L);
    }
    if (items > L.bcap_items) {
        int cap = std::max(items, L.bcap_items * 2);
        dfree(L.d_bitems);
        ECHK(hipMalloc((void **)&L.d_bitems, sizeof(FItem) * cap));
        L.bcap_items = cap;
        destroy_graphs(L);
    }
    return 0;
}

int pool_alloc(int n)
{
    for (auto it = X.free_ranges.begin(); it != X.free_ranges.end(); ++it) {
        if (it->second >= n) {
            const int off = it->first, len = it->second;
            X.free_ranges.erase(it);
            if (len > n) X.free_ranges[off + n] = len - n;
            return off;
        }
    }
    return -1;
}

void pool_free(int off, int n)
{
    auto it = X.free_ranges.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != X.free_ranges.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        X.free_ranges.erase(nx);
    }
    if (it != X.free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            X.free_ranges.erase(it);
        }
    }
}

// Retire DONE / FAILED slots: gather their state on the device, copy once.
int retire()
{
    std::vector<int> done;
    for (int s = 0; s < X.max_slots; ++s)
        if (X.slots[s].busy && (X.h_phase[s] == DONE || X.h_phase[s] == FAILED)) done.push_back(s);
    if (done.empty()) return 0;
    size_t total = 0;
    for (int s : done) total += X.slots[s].n;
    if (ensure_out(total)) return -1;
    size_t k = 0;
    for (int s : done)
        for (int i = 0; i < X.slots[s].n; ++i) X.h_src[k++] = X.slots[s].off + i;
    ECHK(hipMemcpyAsync(X.d_src, X.h_src, sizeof(int) * total, hipMemcpyHostToDevice, X.stream));
    hipLaunchKernelGGL(k_ens_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, X.stream,
                       X.P, X.d_src, X.d_out, (int)total);
    ECHK(hipGetLastError());
    ECHK(hipMemcpyAsync(X.h_out, X.d_out, sizeof(double) * 8 * total, hipMemcpyDeviceToHost,
                        X.stream));
    // final headers of the retired slots: gathered on the device, one download
    for (size_t q = 0; q < done.size(); ++q) X.h_hslots[q] = done[q];
    ECHK(hipMemcpyAsync(X.d_hslots, X.h_hslots, sizeof(int) * done.size(), hipMemcpyHostToDevice,
                        X.stream));
    hipLaunchKernelGGL(k_ens_gather_headers, dim3((unsigned)((done.size() + 255) / 256)),
                       dim3(256), 0, X.stream, X.d_hdr, X.d_hslots, X.d_hstage, (int)done.size());
    ECHK(hipGetLastError());
    ECHK(hipMemcpyAsync(X.h_hstage, X.d_hstage, sizeof(SysHdr) * done.size(),
                        hipMemcpyDeviceToHost, X.stream));
    ECHK(hipStreamSynchronize(X.stream));
    std::vector<SysHdr> hdr(X.h_hstage, X.h_hstage + done.size());
    k = 0;
    for (size_t q = 0; q < done.size(); ++q) {
        SlotInfo &si = X.slots[done[q]];
        Result r;
        r.n = si.n;
        r.x.resize(3 * (size_t)si.n);
        r.v.resize(3 * (size_t)si.n);
        std::vector<double> pot(si.n), pot0(si.n);
        for (int i = 0; i < si.n; ++i) {
            const double *o = X.h_out + 8 * (k + i);
            for (int c = 0; c < 3; ++c) { r.x[3 * i + c] = o[c]; r.v[3 * i + c] = o[3 + c]; }
            pot[i] = o[6];
            pot0[i] = o[7];
        }
        k += si.n;
        r.t = hdr[q].t;
        r.steps = hdr[q].steps;
        r.failed = X.h_phase[done[q]] == FAILED;
        // total energy, kinetic + potential (softened as integrated)
        {
            double kin = 0, pe = 0, kin0 = 0, pe0 = 0;
            for (int i = 0; i < si.n; ++i) {
                const double *vv = &r.v[3 * i];
                kin += si.m[i] * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
                pe += si.m[i] * pot[i];
                const double *v0 = &si.v0[3 * i];
                kin0 += si.m[i] * (v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
                pe0 += si.m[i] * pot0[i];
            }
            r.e1 = 0.5 * kin + 0.5 * pe;
            r.e0 = 0.5 * kin0 + 0.5 * pe0;
        }
        // pair interactions: every particle at the start and at the end, plus
        // each step's particles (all of them in shared mode, the block's in block mode)
        const double psteps = si.mode == 1 ? (double)hdr[q].psteps : (double)r.steps * si.n;
        X.pairs += (psteps + 2.0 * si.n) * (double)(si.n - 1);
        X.results[si.job] = std::move(r);
        X.finished.push_back(si.job);
        X.n_finished++;
        pool_free(si.off, si.n);
        si = SlotInfo();
        X.h_phase[done[q]] = FREE;
    }
    // mark retired slots FREE on the device
    std::vector<SysHdr> zero(done.size());
    for (auto &z : zero) { memset(&z, 0, sizeof z); z.phase = FREE; z.pmin = z.tmin = kInfBits; }
    for (size_t q = 0; q < done.size(); ++q) X.h_hslots[q] = done[q];
    memcpy(X.h_hstage, zero.data(), sizeof(SysHdr) * done.size());
    ECHK(hipMemcpyAsync(X.d_hstage, X.h_hstage, sizeof(SysHdr) * done.size(), hipMemcpyHostToDevice,
                        X.stream));
    ECHK(hipMemcpyAsync(X.d_hslots, X.h_hslots, sizeof(int) * done.size(), hipMemcpyHostToDevice,
                        X.stream));
    hipLaunchKernelGGL(k_ens_set_headers, dim3((unsigned)((done.size() + 255) / 256)), dim3(256),
                       0, X.stream, X.d_hdr, X.d_hstage, X.d_hslots, (int)done.size(), X.d_phase);
    ECHK(hipGetLastError());
    ECHK(hipStreamSynchronize(X.stream));
    X.items_dirty = true;
    return 0;
}

// Admit queued jobs into free slots while the pool has room.
int admit()
{
    if (X.queue.empty()) return 0;
    std::vector<int> new_slots;
    std::vector<SysHdr> new_hdr;
    size_t staged = 0;
    std::vector<std::pair<int, Job>> taken;
    int examined = 0;
    for (auto it = X.queue.begin(); it != X.queue.end() && examined < X.lookahead;) {
        ++examined;
        int slot = -1;
        for (int s = 0; s < X.max_slots; ++s)
            if (!X.slots[s].busy) { slot = s; break; }
        if (slot < 0) break;
        const int off = pool_alloc(it->n);
        if (off < 0) { ++it; continue; }
        SlotInfo &si = X.slots[slot];
        si.busy = true;
        si.job = it->id;
        si.off = off;
        si.n = it->n;
        si.nsplit = nsplit_for(it->n);
        si.lane = lane_for(it->n, it->mode);
        if (si.lane == 0 && X.heavy_lanes == 2 && (slot & 1)) si.lane = 4;
        si.prec = it->prec;
        si.mode = it->mode;
        si.m = it->m;
        si.x0 = it->x;
        si.v0 = it->v;
        si.eps2 = it->eps2;
        si.t_end = it->t_end;
        SysHdr h;
        memset(&h, 0, sizeof h);
        h.t = 0; h.dt = 0; h.t_end = it->t_end; h.eta = it->eta; h.eps2 = it->eps2;
        h.dt_max = std::min(X.dt_max, it->t_end);
        h.pmin = kInfBits;
        h.tmin = kInfBits;
        h.steps = 0;
        h.off = off; h.n = it->n; h.phase = INIT; h.nsplit = si.nsplit; h.prec = it->prec;
        h.lane = si.lane;
        h.mode = it->mode;
        h.jchunk = ((it->n + si.nsplit - 1) / si.nsplit + kTile - 1) / kTile * kTile;
        if (si.nsplit > 1) {
            // split boundaries on whole tiles; recompute the split count
            h.nsplit = (it->n + h.jchunk - 1) / h.jchunk;
            si.nsplit = h.nsplit;
        }
        new_slots.push_back(slot);
        new_hdr.push_back(h);
        staged += it->n;
        taken.emplace_back(slot, std::move(*it));
        it = X.queue.erase(it);
    }
    if (new_slots.empty()) return 0;
    if (ensure_stage(8 * staged)) return -1;
    size_t k = 0;
    for (auto &tj : taken) {
        const Job &j = tj.second;
        const int off = X.slots[tj.first].off;
        for (int i = 0; i < j.n; ++i, ++k) {
            double *r = X.h_stage + 8 * k;
            r[0] = j.m[i];
            for (int c = 0; c < 3; ++c) { r[1 + c] = j.x[3 * i + c]; r[4 + c] = j.v[3 * i + c]; }
            r[7] = (double)(off + i);
        }
    }
    ECHK(hipMemcpyAsync(X.d_stage, X.h_stage, sizeof(double) * 8 * staged, hipMemcpyHostToDevice,
                        X.stream));
    hipLaunchKernelGGL(k_ens_unpack, dim3((unsigned)((staged + 255) / 256)), dim3(256), 0, X.stream,
                       X.P, X.d_stage, (int)staged);
    ECHK(hipGetLastError());
    for (size_t q = 0; q < new_slots.size(); ++q) {
        X.h_hslots[q] = new_slots[q];
        X.h_hstage[q] = new_hdr[q];
    }
    ECHK(hipMemcpyAsync(X.d_hstage, X.h_hstage, sizeof(SysHdr) * new_slots.size(),
                        hipMemcpyHostToDevice, X.stream));
    ECHK(hipMemcpyAsync(X.d_hslots, X.h_hslots, sizeof(int) * new_slots.size(),
                        hipMemcpyHostToDevice, X.stream));
    hipLaunchKernelGGL(k_ens_set_headers, dim3((unsigned)((new_slots.size() + 255) / 256)),
                       dim3(256), 0, X.stream, X.d_hdr, X.d_hstage, X.d_hslots,
                       (int)new_slots.size(), X.d_phase);
    ECHK(hipGetLastError());
    for (size_t q = 0; q < new_slots.size(); ++q) X.h_phase[new_slots[q]] = INIT;
    ECHK(hipStreamSynchronize(X.stream));
    X.items_dirty = true;
    return 0;
}

int rebuild_items()
{
    for (int ln = 2; ln < 4; ++ln) {
        // one entry per system, largest first (the longest workgroups start first)
        Lane &L = X.lanes[ln];
        std::vector<int> sys;
        for (int s = 0; s < X.max_slots; ++s)
            if (X.slots[s].busy && X.slots[s].lane == ln) sys.push_back(s);
        // by group (step mode, then precision: one kernel per group), largest
        // first within each
        auto grp = [](int s) { return 2 * X.slots[s].mode + (X.slots[s].prec == 64 ? 1 : 0); };
        std::stable_sort(sys.begin(), sys.end(), [&](int a, int b) {
            return grp(a) != grp(b) ? grp(a) < grp(b) : X.slots[a].n > X.slots[b].n;
        });
        const int nc = (int)sys.size();
        if (ensure_items(L, std::max(nc, 1), 1)) return -1;
        L.nmax = 0;
        for (int &g : L.ngrp) g = 0;
        for (int q = 0; q < nc; ++q) {
            L.h_chunks[q] = Chunk{sys[q], 0};
            L.nmax = std::max(L.nmax, X.slots[sys[q]].n);
            ++L.ngrp[grp(sys[q])];
        }
        L.n_chunks = nc;
        L.n_items = 0;
        L.h_counts[0] = nc;
        L.h_counts[1] = 0;
        if (nc) ECHK(hipMemcpyAsync(L.d_chunks, L.h_chunks, sizeof(Chunk) * nc,
                                    hipMemcpyHostToDevice, X.stream));
        ECHK(hipMemcpyAsync(L.d_counts, L.h_counts, sizeof(int) * 2, hipMemcpyHostToDevice,
                            X.stream));
    }
    const bool fused_light = X.fused_steps > 0 && X.light_max <= 64;
    for (int ln : {0, 1, 4}) {
        Lane &L = X.lanes[ln];
        // block-step systems take the fused block path, except in the fused
        // light-system kernel, which steps both modes itself
        auto on_bpath = [&](const SlotInfo &si) {
            return si.mode == 1 && X.block_fusion && !(ln == 1 && fused_light);
        };
        int nc = 0, ni = 0, bc = 0, bs = 0, bi = 0;
        for (int s = 0; s < X.max_slots; ++s) {
            const SlotInfo &si = X.slots[s];
            if (!si.busy || si.lane != ln) continue;
            const int ch = (si.n + 63) / 64;
            if (on_bpath(si)) {
                bc += ch;
                bs += 1;
                bi += ch * si.nsplit;
            } else {
                nc += ch;
                ni += ch * si.nsplit;
            }
        }
        if (ensure_items(L, std::max(nc, 1), std::max(ni, 1))) return -1;
        if (ensure_blists(L, std::max(bc, 1), std::max(bs, 1), std::max(bi, 1))) return -1;
        int c = 0, w = 0, q = 0, y = 0;
        for (int s = 0; s < X.max_slots; ++s) {
            const SlotInfo &si = X.slots[s];
            if (!si.busy || si.lane != ln) continue;
            if (on_bpath(si)) {
                L.h_bsys[y++] = s;
                for (int i0 = 0; i0 < si.n; i0 += 64) L.h_schunks[q++] = Chunk{s, i0};
                continue;
            }
            for (int i0 = 0; i0 < si.n; i0 += 64) {
                L.h_chunks[c++] = Chunk{s, i0};
                for (int sp = 0; sp < si.nsplit; ++sp) L.h_items[w++] = FItem{s, i0, sp, 0};
            }
        }
        L.n_chunks = nc;
        L.n_items = ni;
        L.n_schunks = bc;
        L.n_bsys = bs;
        L.n_bitems_max = bi;
        L.h_counts[0] = nc;
        L.h_counts[1] = ni;
        L.h_nsel[0] = bc;
        L.h_nsel[1] = bs;
        if (nc) ECHK(hipMemcpyAsync(L.d_chunks, L.h_chunks, sizeof(Chunk) * nc,
                                    hipMemcpyHostToDevice, X.stream));
        if (ni) ECHK(hipMemcpyAsync(L.d_items, L.h_items, sizeof(FItem) * ni,
                                    hipMemcpyHostToDevice, X.stream));
        ECHK(hipMemcpyAsync(L.d_counts, L.h_counts, sizeof(int) * 2, hipMemcpyHostToDevice,
                            X.stream));
        if (bc) ECHK(hipMemcpyAsync(L.d_schunks, L.h_schunks, sizeof(Chunk) * bc,
                                    hipMemcpyHostToDevice, X.stream));
        if (bs) ECHK(hipMemcpyAsync(L.d_bsys, L.h_bsys, sizeof(int) * bs, hipMemcpyHostToDevice,
                                    X.stream));
        ECHK(hipMemcpyAsync(L.d_nsel, L.h_nsel, sizeof(int) * 2, hipMemcpyHostToDevice,
                            X.stream));
    }
    ECHK(hipStreamSynchronize(X.stream));
    X.items_dirty = false;
    return 0;
}

// Grid bucket: 0 for no work, else the next power of two from 64.
int bucket(int n)
{
    if (n <= 0) return 0;
    int b = 64;
    while (b < n) b *= 2;
    return b;
}

// Waves of the grid-stride block-path kernels: enough to fill the chip,
// never more than the work can use.
constexpr int kBForceWaves = 8192;
constexpr int kBCorrectWaves = 8192;

// The block path's launches, shared by the step graphs and the event-timed
// steps.  sgrid: select chunks (one wave each); ygrid: block-mode systems.
void launch_bselect(const Lane &L, int sgrid)
{
    hipLaunchKernelGGL(k_ens_bselect, dim3(sgrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                       L.d_schunks, L.d_nsel);
}

void launch_bitems(const Lane &L, int ygrid)
{
    const BLists B{L.d_bchunks, L.d_bitems, L.d_bcount};
    hipLaunchKernelGGL(k_ens_bitems, dim3(ygrid / 64), dim3(64), 0, L.stream, X.d_hdr, L.d_bsys,
                       L.d_nsel + 1, B);
}

void launch_bforce(const Lane &L)
{
    hipLaunchKernelGGL(k_ens_bforce, dim3(std::min(kBForceWaves, L.n_bitems_max)), dim3(64), 0,
                       L.stream, X.P, X.d_hdr, L.d_bitems, L.d_bcount + 1);
}

void launch_bcorrect(const Lane &L)
{
    hipLaunchKernelGGL(k_ens_bcorrect, dim3(std::min(kBCorrectWaves, L.n_schunks)), dim3(64), 0,
                       L.stream, X.P, X.d_hdr, L.d_bchunks, L.d_bcount, 1e-30);
}

void enqueue_step(int ln, int cgrid, int igrid, int sgrid, int ygrid)
{
    Lane &L = X.lanes[ln];
    if (cgrid) {
        hipLaunchKernelGGL(k_ens_predict, dim3(cgrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                           L.d_chunks, L.d_counts);
        hipLaunchKernelGGL(k_ens_force, dim3(igrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                           L.d_items, L.d_counts + 1);
        hipLaunchKernelGGL(k_ens_correct, dim3(cgrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                           L.d_chunks, L.d_counts, 1e-30);
    }
    if (sgrid) {
        launch_bselect(L, sgrid);
        launch_bitems(L, ygrid);
        launch_bforce(L);
        launch_bcorrect(L);
    }
    hipLaunchKernelGGL(k_ens_finalize, dim3((X.max_slots + 255) / 256), dim3(256), 0, L.stream,
                       X.d_hdr, X.max_slots, X.d_phase, X.max_steps, 1e-30, ln, L.d_bcount);
}

// One engine step of lane ln launched kernel by kernel, each between two
// events whose elapsed time is added to kernel_ms (option kernel_timing).
int timed_step(int ln)
{
    Lane &L = X.lanes[ln];
    const int cgrid = bucket(L.n_chunks), igrid = bucket(L.n_items);
    const int sgrid = bucket(L.n_schunks), ygrid = bucket(L.n_bsys);
    auto timed = [&](int slot, auto launch) -> int {
        ECHK(hipEventRecord(X.kev[0], L.stream));
        launch();
        ECHK(hipGetLastError());
        ECHK(hipEventRecord(X.kev[1], L.stream));
        ECHK(hipEventSynchronize(X.kev[1]));
        float ms = 0;
        ECHK(hipEventElapsedTime(&ms, X.kev[0], X.kev[1]));
        X.kernel_ms[slot] += ms;
        return 0;
    };
    if (cgrid) {
        if (timed(KS_PREDICT, [&] {
                hipLaunchKernelGGL(k_ens_predict, dim3(cgrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                                   L.d_chunks, L.d_counts);
            }) ||
            timed(KS_FORCE, [&] {
                hipLaunchKernelGGL(k_ens_force, dim3(igrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                                   L.d_items, L.d_counts + 1);
            }) ||
            timed(KS_CORRECT, [&] {
                hipLaunchKernelGGL(k_ens_correct, dim3(cgrid), dim3(64), 0, L.stream, X.P, X.d_hdr,
                                   L.d_chunks, L.d_counts, 1e-30);
            }))
            return -1;
    }
    if (sgrid) {
        if (timed(KS_BSELECT, [&] { launch_bselect(L, sgrid); }) ||
            timed(KS_BITEMS, [&] { launch_bitems(L, ygrid); }) ||
            timed(KS_BFORCE, [&] { launch_bforce(L); }) ||
            timed(KS_BCORRECT, [&] { launch_bcorrect(L); }))
            return -1;
    }
    return timed(KS_FINALIZE, [&] {
        hipLaunchKernelGGL(k_ens_finalize, dim3((X.max_slots + 255) / 256), dim3(256), 0, L.stream,
                           X.d_hdr, X.max_slots, X.d_phase, X.max_steps, 1e-30, ln, L.d_bcount);
    });
}

// A graph of steps_per_graph engine steps of lane ln for its grid buckets.
int get_graph(int ln, hipGraphExec_t &exec)
{
    Lane &L = X.lanes[ln];
    const int cg = bucket(L.n_chunks), ig = bucket(L.n_items);
    const int sg = bucket(L.n_schunks), yg = bucket(L.n_bsys);
    for (auto &g : L.graphs)
        if (g.chunk_grid == cg && g.item_grid == ig && g.sel_grid == sg && g.sys_grid == yg &&
            g.steps == X.steps_per_graph) {
            exec = g.exec;
            return 0;
        }
    hipGraph_t graph;
    ECHK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < X.steps_per_graph; ++s) enqueue_step(ln, cg, ig, sg, yg);
    ECHK(hipStreamEndCapture(L.stream, &graph));
    GraphExec ge;
    ge.chunk_grid = cg;
    ge.item_grid = ig;
    ge.sel_grid = sg;
    ge.sys_grid = yg;
    ge.steps = X.steps_per_graph;
    ECHK(hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    L.graphs.push_back(ge);
    exec = ge.exec;
    return 0;
}

int busy_count()
{
    int b = 0;
    for (auto &s : X.slots) b += s.busy;
    return b;
}

}  // namespace
}  // namespace ens
}  // namespace g6

using namespace g6;
using namespace g6::ens;

namespace {

// Free everything the engine holds (null-safe: also after a failed open).
void release_all()
{
    if (X.stream) (void)hipStreamSynchronize(X.stream);
    for (Lane &L : X.lanes) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        destroy_graphs(L);
        hfree(L.h_chunks); dfree(L.d_chunks); hfree(L.h_items); dfree(L.d_items);
        dfree(L.d_counts); hfree(L.h_counts);
        hfree(L.h_schunks); dfree(L.d_schunks); hfree(L.h_bsys); dfree(L.d_bsys);
        dfree(L.d_bchunks); dfree(L.d_bitems); dfree(L.d_bcount); dfree(L.d_nsel); hfree(L.h_nsel);
        L.chunk_cap = L.item_cap = 0;
        if (L.ev0) (void)hipEventDestroy(L.ev0);
        if (L.ev1) (void)hipEventDestroy(L.ev1);
        if (L.stream) (void)hipStreamDestroy(L.stream);
        L = Lane();
    }
    dfree(X.P.x); dfree(X.P.v); dfree(X.P.a); dfree(X.P.j); dfree(X.P.xp); dfree(X.P.vp);
    dfree(X.P.m); dfree(X.P.pot); dfree(X.P.pot0); dfree(X.P.jp); dfree(X.P.jp64); dfree(X.P.part);
    dfree(X.P.pt); dfree(X.P.pdt); dfree(X.P.act);
    dfree(X.d_hdr); dfree(X.d_phase); hfree(X.h_phase);
    hfree(X.h_hstage); dfree(X.d_hstage); hfree(X.h_hslots); dfree(X.d_hslots);
    hfree(X.h_stage); dfree(X.d_stage); X.stage_cap = 0;
    hfree(X.h_out); dfree(X.d_out); hfree(X.h_src); dfree(X.d_src); X.out_cap = 0;
    for (hipEvent_t &e : X.kev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (X.stream) (void)hipStreamDestroy(X.stream);
    X.stream = nullptr;
    X.slots.clear();
    X.open = false;
}

int open_engine(long long hbm_budget_bytes, int max_systems)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return efail(-1, "g6e_open: no HIP device available");
    if (max_systems < 1 || hbm_budget_bytes <= 0)
        return efail(-1, "g6e_open: bad budget %lld / systems %d", hbm_budget_bytes, max_systems);
    ECHK(hipStreamCreate(&X.stream));
    ECHK(hipEventCreate(&X.kev[0]));
    ECHK(hipEventCreate(&X.kev[1]));
    for (double &k : X.kernel_ms) k = 0;
    // streams are created heavy lanes first: the runtime deals hardware
    // queues to streams in creation order, and the two heavy lanes must not
    // share one to run side by side
    for (int ln : {0, 4, 1, 2, 3}) {
        Lane &L = X.lanes[ln];
        ECHK(hipStreamCreate(&L.stream));
        ECHK(hipEventCreate(&L.ev0));
        ECHK(hipEventCreate(&L.ev1));
        ECHK(hipMalloc((void **)&L.d_counts, sizeof(int) * 2));
        ECHK(hipHostMalloc((void **)&L.h_counts, sizeof(int) * 2, hipHostMallocDefault));
        ECHK(hipMalloc((void **)&L.d_bcount, sizeof(int) * 2));
        ECHK(hipMemset(L.d_bcount, 0, sizeof(int) * 2));
        ECHK(hipMalloc((void **)&L.d_nsel, sizeof(int) * 2));
        ECHK(hipHostMalloc((void **)&L.h_nsel, sizeof(int) * 2, hipHostMallocDefault));
        L.steps = 0;
    }
    // the large workgroup-fused lane needs more than the default dynamic LDS
    {
        const int big = (int)ens_wg_lds(std::max(X.wg_max_n, 1), kWgLargeThreads);
        const int small = (int)ens_wg_lds(std::max(X.wg_small_max, 1), 256);
        ECHK(hipFuncSetAttribute((const void *)k_ens_wg<kWgLargeThreads, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, big));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wg<kWgLargeThreads, 64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, big));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wg<256, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, small));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wg<256, 64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, small));
        const int nbig = std::max(X.wg_max_n, 1), nsmall = std::max(X.wg_small_max, 1);
        ECHK(hipFuncSetAttribute((const void *)k_ens_wgb<kWgbLargeThreads, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ens_wgb_lds(nbig, 32)));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wgb<kWgbLargeThreads, 64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ens_wgb_lds(nbig, 64)));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wgb<256, 32>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ens_wgb_lds(nsmall, 32)));
        ECHK(hipFuncSetAttribute((const void *)k_ens_wgb<256, 64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ens_wgb_lds(nsmall, 64)));
    }
    X.P.max_split = X.max_split;
    const size_t bpp = bytes_per_particle(X.max_split);
    const long long cap = std::max(64LL, hbm_budget_bytes / (long long)bpp);
    if (cap > (1LL << 30)) return efail(-1, "g6e_open: budget too large");
    X.P.cap = (int)cap;
    const size_t c = (size_t)cap;
    ECHK(hipMalloc((void **)&X.P.x, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.v, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.a, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.j, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.xp, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.vp, sizeof(double) * 3 * c));
    ECHK(hipMalloc((void **)&X.P.m, sizeof(double) * c));
    ECHK(hipMalloc((void **)&X.P.pot, sizeof(double) * c));
    ECHK(hipMalloc((void **)&X.P.pot0, sizeof(double) * c));
    ECHK(hipMalloc((void **)&X.P.pt, sizeof(double) * c));
    ECHK(hipMalloc((void **)&X.P.pdt, sizeof(double) * c));
    ECHK(hipMalloc((void **)&X.P.act, sizeof(int) * c));
    ECHK(hipMalloc((void **)&X.P.jp, sizeof(JPred) * c));
    ECHK(hipMalloc((void **)&X.P.jp64, sizeof(JPred64) * c));
    ECHK(hipMalloc((void **)&X.P.part, sizeof(double) * 7 * c * (size_t)X.max_split));
    X.max_slots = max_systems;
    ECHK(hipMalloc((void **)&X.d_hdr, sizeof(SysHdr) * max_systems));
    ECHK(hipMalloc((void **)&X.d_phase, sizeof(int) * max_systems));
    ECHK(hipHostMalloc((void **)&X.h_phase, sizeof(int) * max_systems, hipHostMallocDefault));
    ECHK(hipHostMalloc((void **)&X.h_hstage, sizeof(SysHdr) * max_systems, hipHostMallocDefault));
    ECHK(hipMalloc((void **)&X.d_hstage, sizeof(SysHdr) * max_systems));
    ECHK(hipHostMalloc((void **)&X.h_hslots, sizeof(int) * max_systems, hipHostMallocDefault));
    ECHK(hipMalloc((void **)&X.d_hslots, sizeof(int) * max_systems));
    // all slots FREE
    for (int s = 0; s < max_systems; ++s) {
        memset(&X.h_hstage[s], 0, sizeof(SysHdr));
        X.h_hstage[s].phase = FREE;
        X.h_hstage[s].pmin = X.h_hstage[s].tmin = kInfBits;
        X.h_hslots[s] = s;
        X.h_phase[s] = FREE;
    }
    ECHK(hipMemcpy(X.d_hdr, X.h_hstage, sizeof(SysHdr) * max_systems, hipMemcpyHostToDevice));
    ECHK(hipMemcpy(X.d_phase, X.h_phase, sizeof(int) * max_systems, hipMemcpyHostToDevice));
    X.slots.assign(max_systems, SlotInfo());
    X.free_ranges.clear();
    X.free_ranges[0] = (int)cap;
    X.queue.clear();
    X.results.clear();
    X.finished.clear();
    X.items_dirty = true;
    X.engine_steps = 0;
    X.n_finished = 0;
    X.pairs = 0;
    X.dev_seconds = 0;
    for (double &h : X.host_s) h = 0;
    X.open = true;
    if (rebuild_items()) return -1;
    return 0;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int g6e_open(long long hbm_budget_bytes, int max_systems)
{
    if (X.open) return efail(-1, "g6e_open: already open");
    if (open_engine(hbm_budget_bytes, max_systems)) {
        const std::string err = X.err;
        release_all();   // nothing of a failed open stays allocated
        X.err = err;
        return -1;
    }
    return 0;
}

int g6e_close(void)
{
    if (!X.open) return 0;
    release_all();
    return 0;
}

long long g6e_submit(int n, const double *m, const double *x, const double *v, double t_end,
                     double eta, double eps2)
{
    if (!X.open) return efail(-1, "g6e_submit: engine not open");
    if (n < 1 || n > X.P.cap) return efail(-1, "g6e_submit: n = %d outside [1, %d]", n, X.P.cap);
    if (!m || !x || !v) return efail(-1, "g6e_submit: null m, x or v");
    if (!(t_end > 0) || !(eta > 0) || !(eps2 >= 0))
        return efail(-1, "g6e_submit: bad t_end/eta/eps2");
    Job j;
    j.id = X.next_id++;
    j.n = n;
    j.m.assign(m, m + n);
    j.x.assign(x, x + 3 * (size_t)n);
    j.v.assign(v, v + 3 * (size_t)n);
    j.t_end = t_end;
    j.eta = eta;
    j.eps2 = eps2;
    j.prec = X.precision;
    j.mode = X.block_steps;
    X.queue.push_back(std::move(j));
    return X.next_id - 1;
}

int g6e_run(int steps)
{
    if (!X.open) return efail(-1, "g6e_run: engine not open");
    int done = 0;
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point t0) {
        return std::chrono::duration<double>(clk::now() - t0).count();
    };
    while (true) {
        auto t0 = clk::now();
        if (retire()) return -1;
        X.host_s[0] += since(t0);
        t0 = clk::now();
        if (admit()) return -1;
        X.host_s[1] += since(t0);
        t0 = clk::now();
        if (X.items_dirty && rebuild_items()) return -1;
        X.host_s[2] += since(t0);
        if (busy_count() == 0 || done >= steps) break;
        t0 = clk::now();
        Lane &Lt = X.lanes[1];
        // the heavy lanes: 0, and 4 when heavy systems are split over two
        // streams (one lane's force kernel then overlaps the other's
        // latency-bound select and correct kernels)
        const int hl[2] = {0, 4};
        bool hv[2];
        for (int h = 0; h < 2; ++h)
            hv[h] = X.lanes[hl[h]].n_chunks > 0 || X.lanes[hl[h]].n_schunks > 0;
        const bool heavy = hv[0] || hv[1];
        const bool light = Lt.n_chunks > 0 || Lt.n_schunks > 0;
        const bool med[4] = {false, false, X.lanes[2].n_chunks > 0, X.lanes[3].n_chunks > 0};
        hipGraphExec_t eh[2] = {nullptr, nullptr}, el = nullptr;
        const bool timing = X.kernel_timing != 0;
        for (int h = 0; h < 2; ++h)
            if (hv[h] && !timing && get_graph(hl[h], eh[h])) return -1;
        if (light && !timing && !(X.fused_steps > 0 && X.light_max <= 64) && get_graph(1, el))
            return -1;
        X.host_s[3] += since(t0);
        t0 = clk::now();
        double ms_total = 0;
        for (int h = 0; h < 2; ++h) {
            if (!hv[h]) continue;
            Lane &H = X.lanes[hl[h]];
            ECHK(hipEventRecord(H.ev0, H.stream));
            if (timing) {
                for (int s = 0; s < X.steps_per_graph; ++s)
                    if (timed_step(hl[h])) return -1;
            } else {
                ECHK(hipGraphLaunch(eh[h], H.stream));
            }
            ECHK(hipEventRecord(H.ev1, H.stream));
        }
        for (int ln = 2; ln < 4; ++ln) {
            if (!med[ln]) continue;
            Lane &M = X.lanes[ln];
            ECHK(hipEventRecord(M.ev0, M.stream));
            const size_t lds_wg = ens_wg_lds(M.nmax, ln == 2 ? 256 : kWgLargeThreads);
            const Chunk *lst = M.d_chunks;
            for (int g = 0; g < 4; ++g) {
                const int cnt = M.ngrp[g];
                if (cnt == 0) continue;
                const bool block = g >= 2, f64 = g & 1;
                const size_t lds = block ? ens_wgb_lds(M.nmax, f64 ? 64 : 32) : lds_wg;
                const int iters = block ? X.wgb_steps : X.wg_steps;
#define G6_WG_LAUNCH(KERNEL, NT)                                                               \
    hipLaunchKernelGGL(KERNEL, dim3(cnt), dim3(NT), lds, M.stream, X.P, X.d_hdr, lst, cnt,      \
                       X.d_phase, X.max_steps, 1e-30, iters, M.nmax)
                if (ln == 2) {
                    if (!block && !f64) G6_WG_LAUNCH((k_ens_wg<256, 32>), 256);
                    if (!block && f64) G6_WG_LAUNCH((k_ens_wg<256, 64>), 256);
                    if (block && !f64) G6_WG_LAUNCH((k_ens_wgb<256, 32>), 256);
                    if (block && f64) G6_WG_LAUNCH((k_ens_wgb<256, 64>), 256);
                } else {
                    if (!block && !f64) G6_WG_LAUNCH((k_ens_wg<kWgLargeThreads, 32>), kWgLargeThreads);
                    if (!block && f64) G6_WG_LAUNCH((k_ens_wg<kWgLargeThreads, 64>), kWgLargeThreads);
                    if (block && !f64) G6_WG_LAUNCH((k_ens_wgb<kWgbLargeThreads, 32>), kWgbLargeThreads);
                    if (block && f64) G6_WG_LAUNCH((k_ens_wgb<kWgbLargeThreads, 64>), kWgbLargeThreads);
                }
#undef G6_WG_LAUNCH
                ECHK(hipGetLastError());
                lst += cnt;
            }
            ECHK(hipEventRecord(M.ev1, M.stream));
        }
        auto others_running = [&]() {
            for (int h = 0; h < 2; ++h)
                if (hv[h] && hipStreamQuery(X.lanes[hl[h]].stream) == hipErrorNotReady) return true;
            for (int ln = 2; ln < 4; ++ln)
                if (med[ln] && hipStreamQuery(X.lanes[ln].stream) == hipErrorNotReady) return true;
            return false;
        };
        const bool any_other = heavy || med[2] || med[3];
        if (light) {
            // keep the light lane stepping until the other lanes are done
            const bool fused = X.fused_steps > 0 && X.light_max <= 64;
            int replays = 0;
            do {
                ECHK(hipEventRecord(Lt.ev0, Lt.stream));
                if (fused) {
                    hipLaunchKernelGGL(k_ens_fused, dim3(Lt.n_chunks), dim3(64), 0, Lt.stream, X.P,
                                       X.d_hdr, Lt.d_chunks, Lt.d_counts, X.d_phase, X.max_steps,
                                       1e-30, X.fused_steps);
                    ECHK(hipGetLastError());
                } else if (timing) {
                    for (int s = 0; s < X.steps_per_graph; ++s)
                        if (timed_step(1)) return -1;
                } else {
                    ECHK(hipGraphLaunch(el, Lt.stream));
                }
                ECHK(hipEventRecord(Lt.ev1, Lt.stream));
                ECHK(hipStreamSynchronize(Lt.stream));
                float ms = 0;
                if (timing && fused && hipEventElapsedTime(&ms, Lt.ev0, Lt.ev1) == hipSuccess)
                    X.kernel_ms[KS_FUSED] += ms;
                if (!any_other && hipEventElapsedTime(&ms, Lt.ev0, Lt.ev1) == hipSuccess)
                    ms_total += ms;
                Lt.steps += X.steps_per_graph;
                ++replays;
            } while (any_other && others_running() && replays < 4096);
        }
        double ms_other = 0;
        for (int h = 0; h < 2; ++h) {
            if (!hv[h]) continue;
            Lane &H = X.lanes[hl[h]];
            ECHK(hipStreamSynchronize(H.stream));
            float ms = 0;
            if (hipEventElapsedTime(&ms, H.ev0, H.ev1) == hipSuccess) ms_other = std::max(ms_other, (double)ms);
            H.steps += X.steps_per_graph;
        }
        for (int ln = 2; ln < 4; ++ln) {
            if (!med[ln]) continue;
            Lane &M = X.lanes[ln];
            ECHK(hipStreamSynchronize(M.stream));
            float ms = 0;
            if (hipEventElapsedTime(&ms, M.ev0, M.ev1) == hipSuccess) ms_other = std::max(ms_other, (double)ms);
            M.steps += std::max(X.wg_steps, X.wgb_steps);
        }
        ms_total += ms_other;
        X.host_s[4] += since(t0);
        t0 = clk::now();
        ECHK(hipMemcpyAsync(X.h_phase, X.d_phase, sizeof(int) * X.max_slots, hipMemcpyDeviceToHost,
                            X.stream));
        ECHK(hipStreamSynchronize(X.stream));
        X.host_s[5] += since(t0);
        X.dev_seconds += ms_total * 1e-3;
        // every round counts at least one step, so g6e_run(steps) ends
        const int round_steps = std::max(1, heavy ? X.steps_per_graph
                                            : (med[2] || med[3]) ? std::max(X.wg_steps, X.wgb_steps)
                                                                 : X.steps_per_graph);
        X.engine_steps += round_steps;
        done += round_steps;
    }
    return busy_count() + (int)X.queue.size();
}

int g6e_poll(int max, long long *job_ids)
{
    if (max > 0 && !job_ids) return efail(-1, "g6e_poll: null job_ids");
    int k = 0;
    while (k < max && !X.finished.empty()) {
        job_ids[k++] = X.finished.front();
        X.finished.erase(X.finished.begin());
    }
    return k;
}

int g6e_result(long long job_id, int n, double *x, double *v, double *t_final, long long *steps,
               double *e0, double *e1)
{
    auto it = X.results.find(job_id);
    if (it == X.results.end()) return efail(-1, "g6e_result: job %lld has no result", job_id);
    Result &r = it->second;
    if (n != r.n) return efail(-1, "g6e_result: n = %d, job has %d", n, r.n);
    if (x) memcpy(x, r.x.data(), sizeof(double) * 3 * (size_t)n);
    if (v) memcpy(v, r.v.data(), sizeof(double) * 3 * (size_t)n);
    if (t_final) *t_final = r.t;
    if (steps) *steps = r.steps;
    if (e0) *e0 = r.e0;
    if (e1) *e1 = r.e1;
    const int rc = r.failed ? 1 : 0;
    X.results.erase(it);
    return rc;
}

int g6e_stats(long long *engine_steps, double *pairs, long long *finished, double *device_seconds)
{
    if (engine_steps) *engine_steps = X.engine_steps;
    if (pairs) *pairs = X.pairs;
    if (finished) *finished = X.n_finished;
    if (device_seconds) *device_seconds = X.dev_seconds;
    return 0;
}

int g6e_set_option(const char *key, long long value)
{
    std::string k(key ? key : "");
    if (k == "steps_per_graph") X.steps_per_graph = (int)std::max(1LL, value);
    else if (k == "jsplit_target") X.jsplit_target = (int)std::max(0LL, value);
    else if (k == "max_split") {
        if (X.open) return efail(-1, "max_split is fixed while open");
        X.max_split = (int)std::max(1LL, std::min(64LL, value));
    } else if (k == "lookahead") X.lookahead = (int)std::max(1LL, value);
    else if (k == "dt_max_bits") X.dt_max = std::ldexp(1.0, (int)value);
    else if (k == "max_steps") X.max_steps = std::max(1LL, value);
    else if (k == "block_fusion") {
        // decides the lists a resident block-step system is stepped from
        if (X.open) return efail(-1, "%s is fixed while open", k.c_str());
        X.block_fusion = value ? 1 : 0;
    }
    else if (k == "light_max" || k == "fused_steps" || k == "wg_steps" || k == "wgb_steps" ||
             k == "heavy_lanes") {
        // they decide which lane and kernel each resident system steps in:
        // changed under resident systems, a lane could stop advancing
        if (X.open) return efail(-1, "%s is fixed while open", k.c_str());
        const int v = (int)std::max(0LL, value);
        if (k == "light_max") X.light_max = v;
        else if (k == "fused_steps") X.fused_steps = v;
        else if (k == "wg_steps") X.wg_steps = v;
        else if (k == "heavy_lanes") X.heavy_lanes = v == 2 ? 2 : 1;
        else X.wgb_steps = v;
    }
    else if (k == "wg_small_max" || k == "wg_max_n") {
        // records of 64 B per particle must fit the 160 KB of LDS
        if (X.open) return efail(-1, "%s is fixed while open", k.c_str());
        const int v = (int)std::max(0LL, std::min(2048LL, value));
        if (k == "wg_small_max") X.wg_small_max = v; else X.wg_max_n = v;
    }
    else if (k == "precision") {
        if (value != 32 && value != 64) return efail(-1, "precision must be 32 or 64");
        X.precision = (int)value;
    } else if (k == "kernel_timing") {
        X.kernel_timing = value ? 1 : 0;
    } else if (k == "block_steps") {
        // taken by each job at submission, like the precision
        if (value != 0 && value != 1) return efail(-1, "block_steps must be 0 or 1");
        X.block_steps = (int)value;
    } else return -1;
    return 0;
}

long long g6e_get_option(const char *key)
{
    std::string k(key ? key : "");
    if (k == "steps_per_graph") return X.steps_per_graph;
    if (k == "jsplit_target") return X.jsplit_target;
    if (k == "max_split") return X.max_split;
    if (k == "lookahead") return X.lookahead;
    if (k == "dt_max_bits") return std::ilogb(X.dt_max);
    if (k == "max_steps") return X.max_steps;
    if (k == "precision") return X.precision;
    if (k == "block_steps") return X.block_steps;
    if (k == "block_fusion") return X.block_fusion;
    if (k == "kernel_timing") return X.kernel_timing;
    for (int q = 0; q < KS_N; ++q)
        if (k == std::string("kernel_us_") + kKernelNames[q]) return (long long)(X.kernel_ms[q] * 1e3);
    if (k == "light_max") return X.light_max;
    if (k == "fused_steps") return X.fused_steps;
    if (k == "wg_steps") return X.wg_steps;
    if (k == "wgb_steps") return X.wgb_steps;
    if (k == "heavy_lanes") return X.heavy_lanes;
    if (k == "wg_small_max") return X.wg_small_max;
    if (k == "wg_max_n") return X.wg_max_n;
    if (k == "medium_steps") return X.lanes[2].steps + X.lanes[3].steps;
    static const char *phases[6] = {"host_us_retire", "host_us_admit", "host_us_rebuild",
                                    "host_us_graphs", "host_us_lanes", "host_us_phase"};
    for (int q = 0; q < 6; ++q)
        if (k == phases[q]) return (long long)(X.host_s[q] * 1e6);
    if (k == "light_steps") return X.lanes[1].steps;
    if (k == "heavy_steps") return X.lanes[0].steps + X.lanes[4].steps;
    if (k == "pool_capacity") return X.P.cap;
    if (k == "resident") return busy_count();
    if (k == "queued") return (long long)X.queue.size();
    return -1;
}

const char *g6e_last_error(void) { return X.err.c_str(); }

}  // extern "C"
#pragma GCC visibility pop
