// g6_kernels.hip -- gfx950 kernels of libg6: j scatter (with the i-block
// upload), j predictor, direct-sum force (acc, jerk, pot, nearest neighbour;
// small blocks predict j inside it), split reduction with the FP64
// nearest-neighbour refinement, potential-only and k-nearest-neighbour
// sweeps, and FP64 neighbour lists.
//
// Force pair law: accumulate_acc_and_jerk, src/node/dyn/hdyn/evolve/hdyn_ev.C:2019-2036
//   d = x_j - x_i, w = v_j - v_i, r2 = d.d + eps2
//   pot -= m/r;  acc += m d / r^3;  jerk += m/r^3 (w - 3 (d.w)/r2 d)
// Nearest neighbour: update_nn_coll, hdyn_inline.C:34-53 (smallest d.d,
// strict <, so the lowest address wins ties), excluding the j whose index
// equals index[i] (hdyn_grape6.C:3462-3468).
//
// Precision: positions double-single (hi + lo FP32), pair arithmetic and
// rsqrt FP32, partial sums over j-tiles of 16 to 256 pairs in FP32 folded
// into FP64 accumulators, FP64 across j-splits.  FP64 is also used by the
// predictor and by the exact nearest-neighbour / neighbour-list checks.
#include "g6_internal.h"

#include <climits>

namespace g6 {

// ---------------------------------------------------------------------------
// j-memory update: copy staged records to their addresses.
// ---------------------------------------------------------------------------
// The list holds n addresses, then their n index fields; each record is
// read from the pinned host mirror (mirror[address]) straight into HBM.
// The same launch can copy an i-block of n16 16-byte words from pinned host
// memory to HBM (one launch per block instead of a kernel and a copy).
__global__ void k_scatter(JRaw *__restrict__ jraw, int *__restrict__ jindex,
                          const JRaw *__restrict__ mirror, const int *__restrict__ list, int n,
                          const uint4 *__restrict__ isrc, uint4 *__restrict__ idst, int n16)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n16) idst[k] = isrc[k];
    if (k >= n) return;
    const int a = list[k];
    jraw[a] = mirror[a];
    jindex[a] = list[n + k];
}

hipError_t launch_scatter(hipStream_t s, JRaw *jraw, int *jindex, const JRaw *mirror,
                          const int *list, int n, const void *isrc, void *idst, size_t ibytes)
{
    if (ibytes % 16 || ibytes / 16 > (size_t)INT_MAX) return hipErrorInvalidValue;
    const int n16 = (int)(ibytes / 16);
    const int total = n > n16 ? n : n16;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((total + 255) / 256), dim3(256), 0, s, jraw, jindex,
                       mirror, list, n, static_cast<const uint4 *>(isrc),
                       static_cast<uint4 *>(idst), n16);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// j predictor (FP64, no contraction: rounds exactly as the CPU oracle).
// Horner form of the reference GRAPE simulator, harp3fastsim.c:251-265:
//   xp = ((c*s + a2)*s + v)*s + x,   vp = (c'*s1 + a2)*s2 + v
// with s = ti - tj, s1 = 1.5 s, s2 = 2 s, c = c' = j6 (4th order) or
// c = k18*(0.75 s) + j6, c' = k18*s + j6 (5th order).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split_ds(double x, float &hi, float &lo)
{
    hi = (float)x;
    lo = (float)(x - (double)hi);
}

// One j-particle predicted to ti (the predictor kernel and the fused
// small-block force kernel share it, so both round identically).
template <class Rec>
__device__ __forceinline__ void predict_one(const Rec &r, double ti, int order, double xp[3],
                                            double vp[3])
{
#pragma clang fp contract(off)
    const double s = ti - r.tj;
    const double s1 = 1.5 * s, s2 = 2.0 * s;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double cx = r.j6[k], cv = r.j6[k];
        if (order >= 5) {
            cx = r.k18[k] * (0.75 * s) + r.j6[k];
            cv = r.k18[k] * s + r.j6[k];
        }
        xp[k] = ((cx * s + r.a2[k]) * s + r.v[k]) * s + r.x[k];
        vp[k] = (cv * s1 + r.a2[k]) * s2 + r.v[k];
    }
}

__global__ __launch_bounds__(256) void k_predict(
    const JRaw *__restrict__ jraw, int nj, double ti, int order,
    JPred *__restrict__ jpred, double *__restrict__ xp64,
    double *__restrict__ vp64, int cap)
{
    // m64 = the FP64 masses, right after vp64 ([3][cap] each)
    double *__restrict__ m64 = vp64 + 3 * (size_t)cap;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    double xp[3], vp[3];
    predict_one(jraw[j], ti, order, xp, vp);
    const JRaw &r = jraw[j];
    JPred p;
    split_ds(xp[0], p.xh, p.xl);
    split_ds(xp[1], p.yh, p.yl);
    split_ds(xp[2], p.zh, p.zl);
    p.m = (float)r.m;
    p.vx = (float)vp[0];
    p.vy = (float)vp[1];
    p.vz = (float)vp[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) p.pad[k] = 0.f;
    jpred[j] = p;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xp64[(size_t)k * cap + j] = xp[k];
        vp64[(size_t)k * cap + j] = vp[k];
    }
    m64[j] = r.m;
}

hipError_t launch_predict(hipStream_t s, const JRaw *jraw, int nj, double ti,
                          int order, JPred *jpred, double *xp64, double *vp64,
                          int cap)
{
    if (nj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict, dim3((nj + 255) / 256), dim3(256), 0, s,
                       jraw, nj, ti, order, jpred, xp64, vp64, cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// i-block from j-memory (device-resident sweeps): slot k = address i_first+k.
// ibuf holds 10 arrays of stride ld: xh yh zh xl yl zl vx vy vz excl.
// ---------------------------------------------------------------------------
__global__ void k_iprep_self(const JPred *__restrict__ jpred, int i_first, int ni,
                             float *__restrict__ ibuf, int ld)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ni) return;
    const JPred p = jpred[i_first + k];
    ibuf[0 * ld + k] = p.xh;
    ibuf[1 * ld + k] = p.yh;
    ibuf[2 * ld + k] = p.zh;
    ibuf[3 * ld + k] = p.xl;
    ibuf[4 * ld + k] = p.yl;
    ibuf[5 * ld + k] = p.zl;
    ibuf[6 * ld + k] = p.vx;
    ibuf[7 * ld + k] = p.vy;
    ibuf[8 * ld + k] = p.vz;
    reinterpret_cast<int *>(ibuf)[9 * ld + k] = i_first + k;
}

hipError_t launch_iprep_self(hipStream_t s, const JPred *jpred,
                             const double *, int, int i_first, int ni,
                             float *ibuf, int ld)
{
    if (ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iprep_self, dim3((ni + 255) / 256), dim3(256), 0, s,
                       jpred, i_first, ni, ibuf, ld);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Force kernel.  One wave = 64 i-particles (one per lane) against one j-split;
// j is wave-uniform, so its record comes through scalar loads into SGPRs and
// every VALU instruction does 64 pair interactions.  j is processed in tiles
// of `tile` (64, 128 or 256) j: FP32 partial sums per tile, folded into FP64
// at the tile end.
//
// Nearest neighbour: per lane, the smallest FP32 r^2 of each tile (one
// v_min per pair).  At the tile end the lane keeps the best tile and, as the
// runner-up, the smallest best of the other tiles.  The reduction kernel
// recomputes r^2 over the winning tile with the same FP32 operations, which
// gives the address (first match, so ties resolve to the lowest address as in
// the CPU path) and the tile's own runner-up.  Positive floats compare
// exactly.
// ---------------------------------------------------------------------------
struct Acc32 {
    float ax, ay, az, jx, jy, jz, kx, ky, kz, p;
    float best;
};

// FP32 r^2 of a double-single difference: ((eps2 + dz*dz) + dy*dy) + dx*dx,
// all three steps fused multiply-adds.  Every force variant and the
// reduction kernel's rescan use this exact sequence, so all round
// identically; with eps2 = 0 the first step is exactly dz*dz.
__device__ __forceinline__ float ds_r2(float xjh, float yjh, float zjh, float xjl,
                                       float yjl, float zjl, float xih, float yih,
                                       float zih, float xil, float yil, float zil,
                                       float eps2, float &dx, float &dy, float &dz)
{
#pragma clang fp contract(off)
    dx = (xjh - xih) + (xjl - xil);
    dy = (yjh - yih) + (yjl - yil);
    dz = (zjh - zih) + (zjl - zil);
    float r2 = __builtin_fmaf(dz, dz, eps2);
    r2 = __builtin_fmaf(dy, dy, r2);
    return __builtin_fmaf(dx, dx, r2);
}

// A wave-uniform float copied into a VGPR, for kernels where VALU
// instructions with that SGPR operand measured slower (k_pot, k_force_lds).
__device__ __forceinline__ float to_vgpr(float s)
{
    float v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// The ten fields of a predicted j-particle as the fused small-block kernel
// stages them in LDS (the leading fields of JPred, without its padding).
struct JRec {
    float xh, yh, zh, xl, yl, zl, vx, vy, vz, m;
};

template <bool MASK, bool ZSKIP, class J_T>
__device__ __forceinline__ void pair(const J_T &J, int jj, unsigned rel, float xih,
                                     float yih, float zih, float xil, float yil,
                                     float zil, float vxi, float vyi, float vzi,
                                     float eps2, Acc32 &a)
{
#pragma clang fp contract(off)
    float dx, dy, dz;
    const float r2 = ds_r2(J.xh, J.yh, J.zh, J.xl, J.yl, J.zl, xih, yih, zih, xil,
                           yil, zil, eps2, dx, dy, dz);
    const float dvx = J.vx - vxi;
    const float dvy = J.vy - vyi;
    const float dvz = J.vz - vzi;
    const float rv = __builtin_fmaf(dx, dvx, __builtin_fmaf(dy, dvy, dz * dvz));
    float rinv = __builtin_amdgcn_rsqf(r2);
    float rnn = r2;
    if (ZSKIP) rinv = r2 > 0.0f ? rinv : 0.0f;
    if (MASK) {
        const bool self = (unsigned)jj == rel;
        rinv = self ? 0.0f : rinv;
        rnn = self ? __builtin_inff() : r2;
    }
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
    a.best = __builtin_fminf(a.best, rnn);
}

template <bool MASK, bool ZSKIP, class J_T>
__device__ __forceinline__ void tile_pairs(const J_T *__restrict__ jp, int tn,
                                           unsigned rel, float xih, float yih,
                                           float zih, float xil, float yil,
                                           float zil, float vxi, float vyi,
                                           float vzi, float eps2, Acc32 &a)
{
    int jj = 0;
    for (; jj + 4 <= tn; jj += 4) {
        const J_T *g = jp + jj;
        pair<MASK, ZSKIP>(g[0], jj + 0, rel, xih, yih, zih, xil, yil, zil, vxi, vyi, vzi, eps2, a);
        pair<MASK, ZSKIP>(g[1], jj + 1, rel, xih, yih, zih, xil, yil, zil, vxi, vyi, vzi, eps2, a);
        pair<MASK, ZSKIP>(g[2], jj + 2, rel, xih, yih, zih, xil, yil, zil, vxi, vyi, vzi, eps2, a);
        pair<MASK, ZSKIP>(g[3], jj + 3, rel, xih, yih, zih, xil, yil, zil, vxi, vyi, vzi, eps2, a);
    }
    for (; jj < tn; ++jj)
        pair<MASK, ZSKIP>(jp[jj], jj, rel, xih, yih, zih, xil, yil, zil, vxi, vyi, vzi, eps2, a);
}

// A workgroup holds wpg waves (1..16) on the same 64 i and on consecutive
// j-splits; before writing, the waves fold their partial sums through LDS in
// wave order (a fixed order, so results stay deterministic), so the
// reduction kernel reads nsplit / wpg partials instead of nsplit.  Dynamic
// LDS: double[7][wpg][64], then float best[wpg][64], float second[wpg][64],
// int tile[wpg][64], then float tsum[10][64]: the FP32 sums of each lane's
// winning tile, written by the wave that summed it.
constexpr int kMaxWpg = 16;
constexpr int kTileSums = 10;   // ax ay az jx jy jz kx ky kz pot of one tile
size_t force_fold_lds(int wpg)
{
    return wpg > 1 ? (size_t)wpg * 64 * (7 * 8 + 3 * 4) + 64 * kTileSums * 4 : 0;
}

__device__ __forceinline__ void tile_sums_store(float *p, size_t stride, const Acc32 &a)
{
    p[0 * stride] = a.ax; p[1 * stride] = a.ay; p[2 * stride] = a.az;
    p[3 * stride] = a.jx; p[4 * stride] = a.jy; p[5 * stride] = a.jz;
    p[6 * stride] = a.kx; p[7 * stride] = a.ky; p[8 * stride] = a.kz;
    p[9 * stride] = a.p;
}

__device__ __forceinline__ void tile_sums_load(const float *p, size_t stride, Acc32 &a)
{
    a.ax = p[0 * stride]; a.ay = p[1 * stride]; a.az = p[2 * stride];
    a.jx = p[3 * stride]; a.jy = p[4 * stride]; a.jz = p[5 * stride];
    a.kx = p[6 * stride]; a.ky = p[7 * stride]; a.kz = p[8 * stride];
    a.p = p[9 * stride];
}

// Fused small-block mode (at most kFusedMaxGroups i-groups): each wave
// predicts its own j-split from the raw j records, 64 j at a time (one per
// lane), and stages the FP32 fields in a wave-private LDS area that the pair
// loop reads by broadcast.  The waves of i-group 0 also write the FP64
// prediction (for the reduction's nearest-neighbour rescan, neighbour lists
// and read-back); the other i-groups predict the same j again, from cache.
// No JPred array is written or read, and the predictor launch disappears.
// The pairs are accumulated in the same order as from JPred, so both modes
// give bitwise the same partial sums.
size_t force_fused_lds(int wpg)
{
    const size_t stage = (size_t)wpg * 64 * sizeof(JRec);
    return std::max(stage, force_fold_lds(wpg));   // the fold reuses the staging area
}

// LDS written by some lanes of a wave and read by others of the same wave:
// DS instructions of one wave complete in order, so the fences only keep
// the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool ZSKIP, bool FUSED>
__global__ __launch_bounds__(64 * kMaxWpg) void k_force(
    const JPred *__restrict__ jpred, int nj, IBlock ib, int ni, float eps2,
    int jchunk, int tile, int n_igroups, Partials P, int remap, int wpg, FusedPredict F)
{
    extern __shared__ double fold[];
    int b = blockIdx.x;
    if (remap) {
        // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
        // contiguous run of logical workgroups (same j-splits, so its L2
        // holds only its share of j-memory).
        const int per = gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    const int sgroup = b / n_igroups;
    const int ig = b - sgroup * n_igroups;
    const int lane = threadIdx.x & 63;
    // uniform within the wave: readfirstlane tells the compiler so, which keeps
    // the j records on scalar loads into SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int split = sgroup * wpg + wave;
    const int i = ig * 64 + lane;
    const int ic = i < ni ? i : ni - 1;

    const float xih = ib.xh[ic], yih = ib.yh[ic], zih = ib.zh[ic];
    const float xil = ib.xl[ic], yil = ib.yl[ic], zil = ib.zl[ic];
    const float vxi = ib.vx[ic], vyi = ib.vy[ic], vzi = ib.vz[ic];
    const int excl = i < ni ? ib.excl[ic] : -1;
    // eps2 stays in its SGPR here: a copy in a VGPR measured ~0.5% slower
    // (the j fields, not this constant, set the SGPR-operand cost)

    const int jbeg = split * jchunk;
    const int jend = min(jbeg + jchunk, nj);

    double A0 = 0, A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0;
    float rbest = __builtin_inff(), rsecond = __builtin_inff();
    int rtile = -1;
    // FP32 sums of the best tile so far (the reduction swaps the winning
    // tile's sums for an FP64 recomputation of its pairs)
    Acc32 bt;
    bt.ax = bt.ay = bt.az = bt.jx = bt.jy = bt.jz = bt.kx = bt.ky = bt.kz = bt.p = 0.f;
    bt.best = __builtin_inff();

    for (int t = jbeg; t < jend; t += tile) {
        const int tn = min(tile, jend - t);
        Acc32 a;
        a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.kx = a.ky = a.kz = a.p = 0.f;
        a.best = __builtin_inff();
        const unsigned rel = (unsigned)(excl - t);
        if (FUSED) {
            JRec *stage = reinterpret_cast<JRec *>(fold) + wave * 64;
            for (int s0 = 0; s0 < tn; s0 += 64) {
                const int sn = min(64, tn - s0);
                wave_lds_sync();   // the previous chunk has been read
                if (lane < sn) {
                    const int j = t + s0 + lane;
                    const JRaw &r = F.jraw[j];
                    double xp[3], vp[3];
                    predict_one(r, F.ti, F.order, xp, vp);
                    if (ig == 0) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            F.xp64[(size_t)k * F.cap + j] = xp[k];
                            F.vp64[(size_t)k * F.cap + j] = vp[k];
                        }
                        F.vp64[3 * (size_t)F.cap + j] = r.m;   // m64
                    }
                    JRec q;
                    split_ds(xp[0], q.xh, q.xl);
                    split_ds(xp[1], q.yh, q.yl);
                    split_ds(xp[2], q.zh, q.zl);
                    q.vx = (float)vp[0];
                    q.vy = (float)vp[1];
                    q.vz = (float)vp[2];
                    q.m = (float)r.m;
                    stage[lane] = q;
                }
                wave_lds_sync();
                const unsigned rs = rel - (unsigned)s0;
                if (__ballot(rs < (unsigned)sn) == 0)
                    tile_pairs<false, ZSKIP>(stage, sn, rs, xih, yih, zih, xil, yil, zil, vxi,
                                             vyi, vzi, eps2, a);
                else
                    tile_pairs<true, ZSKIP>(stage, sn, rs, xih, yih, zih, xil, yil, zil, vxi,
                                            vyi, vzi, eps2, a);
            }
        } else {
            const bool mine = rel < (unsigned)tn;
            if (__ballot(mine) == 0)
                tile_pairs<false, ZSKIP>(jpred + t, tn, rel, xih, yih, zih, xil, yil, zil,
                                         vxi, vyi, vzi, eps2, a);
            else
                tile_pairs<true, ZSKIP>(jpred + t, tn, rel, xih, yih, zih, xil, yil, zil,
                                        vxi, vyi, vzi, eps2, a);
        }
        A0 += (double)a.ax;
        A1 += (double)a.ay;
        A2 += (double)a.az;
        A3 += (double)a.jx - 3.0 * (double)a.kx;
        A4 += (double)a.jy - 3.0 * (double)a.ky;
        A5 += (double)a.jz - 3.0 * (double)a.kz;
        A6 += (double)a.p;
        // runner-up: the smallest best among the other tiles
        if (a.best < rbest) {
            rsecond = __builtin_fminf(rsecond, rbest);
            rbest = a.best;
            rtile = t;
            bt = a;
        } else {
            rsecond = __builtin_fminf(rsecond, a.best);
        }
    }

    if (wpg > 1) {
        // in fused mode the fold area is the staging area: every wave must
        // have left its j loop first
        if (FUSED) __syncthreads();
        const int stride = wpg * 64;
        const int slot = wave * 64 + lane;
        float *fb = reinterpret_cast<float *>(fold + 7 * stride);
        float *fs = fb + stride;
        int *ft = reinterpret_cast<int *>(fs + stride);
        fold[0 * stride + slot] = A0;
        fold[1 * stride + slot] = A1;
        fold[2 * stride + slot] = A2;
        fold[3 * stride + slot] = A3;
        fold[4 * stride + slot] = A4;
        fold[5 * stride + slot] = A5;
        fold[6 * stride + slot] = A6;
        fb[slot] = rbest;
        fs[slot] = rsecond;
        ft[slot] = rtile;
        __syncthreads();
        // the wave whose split holds this lane's winning tile, by the fold's
        // rule below (first strict minimum in wave order), hands its tile
        // sums to wave 0
        float *fts = reinterpret_cast<float *>(ft + stride);   // [kTileSums][64]
        int win = 0;
        {
            float wb = fb[lane];
            for (int w = 1; w < wpg; ++w) {
                const float bw = fb[w * 64 + lane];
                if (bw < wb) { wb = bw; win = w; }
            }
        }
        if (win == wave && wave != 0) tile_sums_store(fts + lane, 64, bt);
        __syncthreads();
        if (wave != 0) return;
        if (win != 0) tile_sums_load(fts + lane, 64, bt);
        for (int w = 1; w < wpg; ++w) {
            const int o = w * 64 + lane;
            A0 += fold[0 * stride + o];
            A1 += fold[1 * stride + o];
            A2 += fold[2 * stride + o];
            A3 += fold[3 * stride + o];
            A4 += fold[4 * stride + o];
            A5 += fold[5 * stride + o];
            A6 += fold[6 * stride + o];
            // later splits hold higher addresses: they win only strictly
            const float bw = fb[o];
            if (bw < rbest) {
                rsecond = __builtin_fminf(rbest, fs[o]);
                rbest = bw;
                rtile = ft[o];
            } else {
                rsecond = __builtin_fminf(rsecond, bw);
            }
        }
    }
    if (i < ni) {
        const size_t o = (size_t)sgroup * P.ld + i;
        P.c[0][o] = A0;
        P.c[1][o] = A1;
        P.c[2][o] = A2;
        P.c[3][o] = A3;
        P.c[4][o] = A4;
        P.c[5][o] = A5;
        P.c[6][o] = A6;
        P.best[o] = rbest;
        P.btile[o] = rtile;
        P.second[o] = rsecond;
        tile_sums_store(P.tsum + o, P.tstride, bt);
    }
}

// ---------------------------------------------------------------------------
// LDS-staged force kernel.  A workgroup of 256 lanes stages 256-j tiles of
// predicted j-particles in LDS; every lane reads each j by broadcast into
// VGPRs and accumulates it onto two i-particles (i and i + 256 of the
// workgroup's 512).  With j in VGPRs the pair arithmetic is VOP2 FP32
// add/sub/mul/fmac with vector operands only, which gfx950 issues at twice
// the rate of forms that read an SGPR.  Two i per lane halve the LDS bytes per
// pair interaction (40 B per j per lane, shared by two pairs).
// Per-tile FP32 sums, nearest-neighbour tracking and partial output are the
// same as k_force.
// ---------------------------------------------------------------------------
struct IPos {
    float xh, yh, zh, xl, yl, zl, vx, vy, vz;
};

template <bool MASK, bool EPS, bool ZSKIP>
__device__ __forceinline__ void pair_v(float jxh, float jyh, float jzh, float jxl,
                                       float jyl, float jzl, float jvx, float jvy,
                                       float jvz, float jm, int jj, unsigned rel,
                                       const IPos &I, float eps2v, Acc32 &a)
{
#pragma clang fp contract(off)
    float dx, dy, dz;
    const float r2 = ds_r2(jxh, jyh, jzh, jxl, jyl, jzl, I.xh, I.yh, I.zh, I.xl, I.yl, I.zl,
                           EPS ? eps2v : 0.0f, dx, dy, dz);
    const float dvx = jvx - I.vx;
    const float dvy = jvy - I.vy;
    const float dvz = jvz - I.vz;
    float rv = dz * dvz;
    rv = __builtin_fmaf(dy, dvy, rv);
    rv = __builtin_fmaf(dx, dvx, rv);
    float rinv = __builtin_amdgcn_rsqf(r2);
    float rnn = r2;
    if (ZSKIP) rinv = r2 > 0.0f ? rinv : 0.0f;
    if (MASK) {
        const bool self = (unsigned)jj == rel;
        rinv = self ? 0.0f : rinv;
        rnn = self ? __builtin_inff() : r2;
    }
    const float rinv2 = rinv * rinv;
    const float mr = jm * rinv;
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
    a.best = __builtin_fminf(a.best, rnn);
}

constexpr int kLdsWG = 256;   // lanes per workgroup of the LDS kernel
constexpr int kLdsI = 2;      // i-particles per lane

template <bool M0, bool M1, bool EPS, bool ZSKIP>
__device__ __forceinline__ void tile_lds(const float4 *__restrict__ sj, int tn,
                                         unsigned rel0, unsigned rel1, const IPos &I0,
                                         const IPos &I1, float eps2v, Acc32 &a0,
                                         Acc32 &a1)
{
#pragma unroll 2
    for (int jj = 0; jj < tn; ++jj) {
        const float4 A = sj[3 * jj + 0];   // xh yh zh xl
        const float4 B = sj[3 * jj + 1];   // yl zl vx vy
        const float4 C = sj[3 * jj + 2];   // vz m - -
        pair_v<M0, EPS, ZSKIP>(A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, C.x, C.y, jj, rel0,
                               I0, eps2v, a0);
        pair_v<M1, EPS, ZSKIP>(A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, C.x, C.y, jj, rel1,
                               I1, eps2v, a1);
    }
}

__device__ __forceinline__ void acc_reset(Acc32 &a)
{
    a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.kx = a.ky = a.kz = a.p = 0.f;
    a.best = __builtin_inff();
}

struct Acc64 {
    double c[7];
    float best, second;
    int tile;
    Acc32 bt;   // FP32 sums of the best tile
};

__device__ __forceinline__ void acc_fold(Acc64 &A, const Acc32 &a, int t)
{
    A.c[0] += (double)a.ax;
    A.c[1] += (double)a.ay;
    A.c[2] += (double)a.az;
    A.c[3] += (double)a.jx - 3.0 * (double)a.kx;
    A.c[4] += (double)a.jy - 3.0 * (double)a.ky;
    A.c[5] += (double)a.jz - 3.0 * (double)a.kz;
    A.c[6] += (double)a.p;
    if (a.best < A.best) {
        A.second = __builtin_fminf(A.second, A.best);
        A.best = a.best;
        A.tile = t;
        A.bt = a;
    } else {
        A.second = __builtin_fminf(A.second, a.best);
    }
}

__device__ __forceinline__ IPos load_i(const IBlock &ib, int ic)
{
    IPos I;
    I.xh = ib.xh[ic]; I.yh = ib.yh[ic]; I.zh = ib.zh[ic];
    I.xl = ib.xl[ic]; I.yl = ib.yl[ic]; I.zl = ib.zl[ic];
    I.vx = ib.vx[ic]; I.vy = ib.vy[ic]; I.vz = ib.vz[ic];
    return I;
}

template <bool EPS, bool ZSKIP>
__global__ __launch_bounds__(256) void k_force_lds(
    const JPred *__restrict__ jpred, int nj, IBlock ib, int ni, float eps2,
    int jchunk, int n_igroups, Partials P, int remap)
{
    __shared__ float4 sj[3 * kTile];
    int b = blockIdx.x;
    if (remap) {
        const int per = gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    const int split = b / n_igroups;
    const int ig = b - split * n_igroups;
    const int tid = threadIdx.x;
    const int i0 = ig * (kLdsWG * kLdsI) + tid;
    const int i1 = i0 + kLdsWG;
    const IPos I0 = load_i(ib, i0 < ni ? i0 : ni - 1);
    const IPos I1 = load_i(ib, i1 < ni ? i1 : ni - 1);
    const int ex0 = i0 < ni ? ib.excl[i0] : -1;
    const int ex1 = i1 < ni ? ib.excl[i1] : -1;
    const float eps2v = to_vgpr(eps2);

    const int jbeg = split * jchunk;
    const int jend = min(jbeg + jchunk, nj);

    Acc64 A0, A1;
#pragma unroll
    for (int c = 0; c < 7; ++c) A0.c[c] = A1.c[c] = 0.0;
    A0.best = A0.second = A1.best = A1.second = __builtin_inff();
    A0.tile = A1.tile = -1;
    acc_reset(A0.bt);
    acc_reset(A1.bt);

    const float4 *g4 = reinterpret_cast<const float4 *>(jpred);
    for (int t = jbeg; t < jend; t += kTile) {
        const int tn = min(kTile, jend - t);
        __syncthreads();                       // previous tile fully read
        if (tid < tn) {
            const float4 p0 = g4[4 * (size_t)(t + tid) + 0];   // xh yh zh xl
            const float4 p1 = g4[4 * (size_t)(t + tid) + 1];   // yl zl vx vy
            const float4 p2 = g4[4 * (size_t)(t + tid) + 2];   // vz m pad pad
            sj[3 * tid + 0] = p0;
            sj[3 * tid + 1] = p1;
            sj[3 * tid + 2] = p2;
        }
        __syncthreads();
        Acc32 a0, a1;
        acc_reset(a0);
        acc_reset(a1);
        const unsigned rel0 = (unsigned)(ex0 - t);
        const unsigned rel1 = (unsigned)(ex1 - t);
        const bool m0 = __ballot(rel0 < (unsigned)tn) != 0;
        const bool m1 = __ballot(rel1 < (unsigned)tn) != 0;
        if (!m0 && !m1)
            tile_lds<false, false, EPS, ZSKIP>(sj, tn, rel0, rel1, I0, I1, eps2v, a0, a1);
        else if (m0 && !m1)
            tile_lds<true, false, EPS, ZSKIP>(sj, tn, rel0, rel1, I0, I1, eps2v, a0, a1);
        else if (!m0 && m1)
            tile_lds<false, true, EPS, ZSKIP>(sj, tn, rel0, rel1, I0, I1, eps2v, a0, a1);
        else
            tile_lds<true, true, EPS, ZSKIP>(sj, tn, rel0, rel1, I0, I1, eps2v, a0, a1);
        acc_fold(A0, a0, t);
        acc_fold(A1, a1, t);
    }

    const size_t o0 = (size_t)split * P.ld + i0;
    if (i0 < ni) {
#pragma unroll
        for (int c = 0; c < 7; ++c) P.c[c][o0] = A0.c[c];
        P.best[o0] = A0.best;
        P.btile[o0] = A0.tile;
        P.second[o0] = A0.second;
        tile_sums_store(P.tsum + o0, P.tstride, A0.bt);
    }
    if (i1 < ni) {
        const size_t o1 = o0 + kLdsWG;
#pragma unroll
        for (int c = 0; c < 7; ++c) P.c[c][o1] = A1.c[c];
        P.best[o1] = A1.best;
        P.btile[o1] = A1.tile;
        P.second[o1] = A1.second;
        tile_sums_store(P.tsum + o1, P.tstride, A1.bt);
    }
}

int force_igroup(int variant) { return variant == 1 ? kLdsWG * kLdsI : 64; }

int force_partials(int variant, int nsplit, int wpg)
{
    if (variant == 1 || wpg <= 1) return nsplit;
    return (nsplit + wpg - 1) / wpg;
}

hipError_t launch_force(hipStream_t s, const JPred *jpred, int nj,
                        const IBlock &ib, int ni, float eps2, int nsplit,
                        int jchunk, int tile, const Partials &p, bool xcd_remap, int variant,
                        bool zskip, int wpg, const FusedPredict *fused)
{
    if (ni <= 0) return hipSuccess;
    if (variant == 1) wpg = 1;
    if (wpg < 1 || wpg > kMaxWpg || (wpg & (wpg - 1))) return hipErrorInvalidValue;
    const int n_igroups = (ni + force_igroup(variant) - 1) / force_igroup(variant);
    // fused prediction: the scalar-load kernel with few i-groups; the FP64
    // prediction of each j is written by exactly one lane (i-group 0)
    if (fused && (variant != 0 || n_igroups > kFusedMaxGroups)) return hipErrorInvalidValue;
    const int nwg = n_igroups * force_partials(variant, nsplit, wpg);
    const int remap = (xcd_remap && nwg % 8 == 0) ? 1 : 0;
    // r^2 >= eps2 > 0 when eps2 is a normal float, so the zero-r^2 rule only
    // needs its per-pair select without softening
    zskip = zskip && !(eps2 >= __FLT_MIN__);
    if (variant == 1) {
        if (zskip)
            hipLaunchKernelGGL((k_force_lds<false, true>), dim3(nwg), dim3(kLdsWG), 0, s, jpred,
                               nj, ib, ni, 0.0f, jchunk, n_igroups, p, remap);
        else if (eps2 != 0.0f)
            hipLaunchKernelGGL((k_force_lds<true, false>), dim3(nwg), dim3(kLdsWG), 0, s, jpred,
                               nj, ib, ni, eps2, jchunk, n_igroups, p, remap);
        else
            hipLaunchKernelGGL((k_force_lds<false, false>), dim3(nwg), dim3(kLdsWG), 0, s, jpred,
                               nj, ib, ni, eps2, jchunk, n_igroups, p, remap);
    } else if (fused) {
        const FusedPredict F = *fused;
        const size_t lds = force_fused_lds(wpg);
        if (zskip)
            hipLaunchKernelGGL((k_force<true, true>), dim3(nwg), dim3(64 * wpg), lds, s, jpred,
                               nj, ib, ni, 0.0f, jchunk, tile, n_igroups, p, remap, wpg, F);
        else
            hipLaunchKernelGGL((k_force<false, true>), dim3(nwg), dim3(64 * wpg), lds, s, jpred,
                               nj, ib, ni, eps2, jchunk, tile, n_igroups, p, remap, wpg, F);
    } else if (zskip) {
        hipLaunchKernelGGL((k_force<true, false>), dim3(nwg), dim3(64 * wpg), force_fold_lds(wpg),
                           s, jpred, nj, ib, ni, 0.0f, jchunk, tile, n_igroups, p, remap, wpg,
                           FusedPredict{});
    } else {
        hipLaunchKernelGGL((k_force<false, false>), dim3(nwg), dim3(64 * wpg),
                           force_fold_lds(wpg), s, jpred, nj, ib, ni, eps2, jchunk, tile,
                           n_igroups, p, remap, wpg, FusedPredict{});
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reduction over j-splits into the caller's layout.  A group of L = 2^LOG2L
// lanes handles one i (64 / L i per wave): few splits use small groups, many
// splits (small i-blocks) a whole wave.
//
// Lane l of a group sums splits l, l + L, ... in order; a butterfly of xor
// shuffles inside the group then adds across lanes.  Every level adds the
// same two values in both partner lanes (IEEE addition commutes), so all
// lanes of the group end with one result and the sum depends only on the
// split geometry: deterministic.
//
// Nearest neighbour: the (FP32 r^2, tile) minimum over splits, lowest tile
// on ties, and the runner-up over everything else.  The address inside the
// winning tile is found by recomputing FP32 r^2 exactly as the force kernel
// did (first match in ascending address).  When the runner-up lies within
// kNNBand of the best, the group rescans in FP64 the j of every split whose
// own best lies within the band (d^2 evaluated as the FP64 CPU path,
// lexicographic (d^2, address) minimum) and flags the i as refined; when
// nothing matched it rescans all j.
//
// Nearest tile in FP64: the winning tile holds the closest pair, which
// usually dominates acc and jerk, and its FP32 sums carry that pair's
// rounding and the rounding of every later pair added at its magnitude.  The
// same scan recomputes the tile's pairs in FP64 (the FP64 pair law of the
// CPU path, hdyn_ev.C:2019-2036) and the result is
//   sum over splits - (the tile's FP32 sums as folded) + (the FP64 pairs),
// so the closest pair no longer limits the relative error.
//
// Every branch below depends only on values that are equal across the
// group, and shuffles stay inside the group, so groups of one wave may
// diverge freely.
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

constexpr int kReduceWaves = 4;   // waves per reduction workgroup

// One pair of the FP64 pair law (hdyn_ev.C:2019-2036), added to F[0..6] =
// acc, jerk, pot.  1/r comes from the hardware FP64 reciprocal square root
// (relative error ~2^-23) refined by two Newton steps, y <- y + y/2 (1 - r2
// y^2), each squaring the error: ~2^-46, then below an FP64 rounding.  That
// is a quarter of the instructions of a correctly rounded divide and square
// root, and agrees with them to a few FP64 ulps.
__device__ __forceinline__ void pair64_r2(double dx, double dy, double dz, double du,
                                          double dv, double dw, double m, double r2,
                                          double F[7])
{
    double y = __builtin_amdgcn_rsq(r2);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = __builtin_fma(-(r2 * y), y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    const double r2inv = y * y;
    const double a3 = -3.0 * (dx * du + dy * dv + dz * dw) * r2inv;
    const double mrinv = m * y;
    const double mr3inv = mrinv * r2inv;
    F[0] += mr3inv * dx;
    F[1] += mr3inv * dy;
    F[2] += mr3inv * dz;
    F[3] += mr3inv * (du + a3 * dx);
    F[4] += mr3inv * (dv + a3 * dy);
    F[5] += mr3inv * (dw + a3 * dz);
    F[6] -= mrinv;
}

__device__ __forceinline__ void pair64(double dx, double dy, double dz, double du, double dv,
                                       double dw, double m, double eps2, double F[7])
{
    pair64_r2(dx, dy, dz, du, dv, dw, m, dx * dx + dy * dy + dz * dz + eps2, F);
}

template <int LOG2L>
__global__ __launch_bounds__(64 * kReduceWaves) void k_reduce(
    Partials P, int nsplit, int ni, int nj, IBlock ib,
    float eps2, int tile_len, const double *__restrict__ xp64, int cap,
    const int *__restrict__ jindex, Results R, NearTile NT, unsigned *ticket,
    unsigned long long *done_word, unsigned long long seq)
{
#pragma clang fp contract(off)
    constexpr int L = 1 << LOG2L;
    const int lane = threadIdx.x & 63;
    const int l = lane & (L - 1);
    const int ig = (blockIdx.x * kReduceWaves + (threadIdx.x >> 6)) * (64 / L) + (lane >> LOG2L);
    // groups past the block's end redo its last i and write nothing, so that
    // every thread reaches the completion ticket below
    const bool live = ig < ni;
    const int i = live ? ig : ni - 1;

    double A[7] = {0, 0, 0, 0, 0, 0, 0};
    float best = __builtin_inff(), second = __builtin_inff();
    int tile = INT_MAX;
    // splits l, l + L, ... in order; loads issued four splits at a time
    constexpr int U = 4;
    for (int s0 = l; s0 < nsplit; s0 += U * L) {
        double v[U][7];
        float vb[U], vs[U];
        int vt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = min(s0 + u * L, nsplit - 1);
            const size_t o = (size_t)s * P.ld + i;
#pragma unroll
            for (int c = 0; c < 7; ++c) v[u][c] = P.c[c][o];
            vb[u] = P.best[o];
            vs[u] = P.second[o];
            vt[u] = P.btile[o];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s0 + u * L >= nsplit) break;
#pragma unroll
            for (int c = 0; c < 7; ++c) A[c] += v[u][c];
            const float b = vb[u];
            if (b < best) {
                second = __builtin_fminf(best, vs[u]);
                best = b;
                tile = vt[u];
            } else {
                second = __builtin_fminf(second, b);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) A[c] = group_sum<L>(A[c]);
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const float os = __shfl_xor(second, off);
        const int ot = __shfl_xor(tile, off);
        const bool take = ob < best || (ob == best && ot < tile);
        const float loser = take ? best : ob;
        second = __builtin_fminf(__builtin_fminf(second, os), loser);
        best = take ? ob : best;
        tile = take ? ot : tile;
    }

    const int excl = ib.excl[i];
    int nn = 0;
    bool ambiguous = !(best < __builtin_inff()) || tile < 0 || tile == INT_MAX;
    bool narrow = false;   // ambiguous only through the band test
    if (!ambiguous) {
        const float xih = ib.xh[i], yih = ib.yh[i], zih = ib.zh[i];
        const float xil = ib.xl[i], yil = ib.yl[i], zil = ib.zl[i];
        const double xi = ib.x64[i], yi = ib.y64[i], zi = ib.z64[i];
        const double ui = ib.vx64[i], vi = ib.vy64[i], wi = ib.vz64[i];
        const int tend = min(tile + tile_len, nj);
        const uint32_t bb = __float_as_uint(best);
        const int gshift = lane & ~(L - 1);
        // the tile's FP32 sums as the force kernel kept them, loaded before
        // the scan so their latency overlaps it
        const size_t o = (size_t)((tile / NT.jchunk) / NT.wpg) * P.ld + i;
        Acc32 t;
        tile_sums_load(P.tsum + o, P.tstride, t);
        int found = INT_MAX;
        // the whole tile: first address matching best, the two smallest r^2
        // (as a multiset) for the tile's own runner-up, and its pairs in FP64
        float m1 = __builtin_inff(), m2 = __builtin_inff();
        double F[7] = {0, 0, 0, 0, 0, 0, 0};
        // L j per pass; the loads of kScanPasses passes are issued together
        // (a 256-j tile at L = 32 is 8 passes, each a memory round trip when
        // taken one at a time), then the passes are evaluated in address
        // order, so sums and the first match are those of one pass at a time
        constexpr int kScanPasses = 4;
        for (int j00 = tile; j00 < tend; j00 += kScanPasses * L) {
            double q[kScanPasses][7];
#pragma unroll
            for (int u = 0; u < kScanPasses; ++u) {
                const int j = min(j00 + u * L + l, tend - 1);   // in range; unused past tend
                q[u][0] = xp64[j];
                q[u][1] = xp64[(size_t)cap + j];
                q[u][2] = xp64[2 * (size_t)cap + j];
                q[u][3] = NT.vp64[j];
                q[u][4] = NT.vp64[(size_t)cap + j];
                q[u][5] = NT.vp64[2 * (size_t)cap + j];
                q[u][6] = NT.m64[j];
            }
#pragma unroll
            for (int u = 0; u < kScanPasses; ++u) {
                const int j0 = j00 + u * L;
                const int j = j0 + l;
                bool hit = false;
                if (j < tend && j != excl) {
                    // the double-single words the force kernel used, split
                    // again from the FP64 prediction they were made from
                    const double xj = q[u][0], yj = q[u][1], zj = q[u][2];
                    float xjh, xjl, yjh, yjl, zjh, zjl;
                    split_ds(xj, xjh, xjl);
                    split_ds(yj, yjh, yjl);
                    split_ds(zj, zjh, zjl);
                    float dx, dy, dz;
                    const float r2 = ds_r2(xjh, yjh, zjh, xjl, yjl, zjl, xih, yih, zih, xil,
                                           yil, zil, eps2, dx, dy, dz);
                    hit = __float_as_uint(r2) == bb;
                    m2 = __builtin_amdgcn_fmed3f(m1, m2, r2);
                    m1 = __builtin_fminf(m1, r2);
                    if (!(NT.zskip && r2 == 0.0f))
                        pair64(xj - xi, yj - yi, zj - zi, q[u][3] - ui, q[u][4] - vi,
                               q[u][5] - wi, q[u][6], NT.eps2d, F);
                }
                unsigned long long m = __ballot(hit) >> gshift;
                if (L < 64) m &= (1ull << L) - 1ull;
                if (m && found == INT_MAX) found = j0 + __ffsll((long long)m) - 1;
            }
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            const float o1 = __shfl_xor(m1, off);
            const float o2 = __shfl_xor(m2, off);
            m2 = __builtin_fminf(__builtin_fmaxf(m1, o1), __builtin_fminf(m2, o2));
            m1 = __builtin_fminf(m1, o1);
        }
        // swap the tile's FP32 sums (converted and combined exactly as the
        // force kernel folded them) for its FP64 pairs
        const double c32[7] = {(double)t.ax, (double)t.ay, (double)t.az,
                               (double)t.jx - 3.0 * (double)t.kx,
                               (double)t.jy - 3.0 * (double)t.ky,
                               (double)t.jz - 3.0 * (double)t.kz, (double)t.p};
#pragma unroll
        for (int c = 0; c < 7; ++c) A[c] = (A[c] - c32[c]) + group_sum<L>(F[c]);
        second = __builtin_fminf(second, m2);
        if (found == INT_MAX) ambiguous = true;
        else nn = jindex[found];
        if (!(second > best * (1.0f + kNNBand))) {
            narrow = !ambiguous;
            ambiguous = true;
        }
    }
    if (l == 0 && live) {
        R.acc[3 * i + 0] = A[0];
        R.acc[3 * i + 1] = A[1];
        R.acc[3 * i + 2] = A[2];
        R.jerk[3 * i + 0] = A[3];
        R.jerk[3 * i + 1] = A[4];
        R.jerk[3 * i + 2] = A[5];
        R.pot[i] = A[6];
    }
    const bool refine = ambiguous && nj > 1;
    if (refine) {
        const double xi = ib.x64[i], yi = ib.y64[i], zi = ib.z64[i];
        double bd = __builtin_inf();
        int ba = INT_MAX;
        // FP32 r^2 is within ~2^-21 of the FP64 d^2 (+ eps2), so the FP64
        // nearest has FP32 r^2 <= best (1 + kNNBand): when the ambiguity
        // came only from the band test it lies in a partial row (j range
        // [s jchunk wpg, (s + 1) jchunk wpg)) whose own best is within that
        // bound, and only those rows are rescanned, in ascending address as
        // the whole-range scan would.  Otherwise every j is rescanned.
        const float thr = best * (1.0f + kNNBand);
        const int span = NT.jchunk * NT.wpg;
        for (int s = 0; s < (narrow ? nsplit : 1); ++s) {
            int jb = 0, je = nj;
            if (narrow) {
                if (!(P.best[(size_t)s * P.ld + i] <= thr)) continue;
                jb = s * span;
                je = min(nj, jb + span);
            }
            for (int j = jb + l; j < je; j += L) {
                if (j == excl) continue;
                const double dx = xp64[j] - xi;
                const double dy = xp64[(size_t)cap + j] - yi;
                const double dz = xp64[2 * (size_t)cap + j] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < bd) { bd = d2; ba = j; }
            }
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            const double od = __shfl_xor(bd, off);
            const int oa = __shfl_xor(ba, off);
            if (od < bd || (od == bd && oa < ba)) { bd = od; ba = oa; }
        }
        if (ba != INT_MAX) nn = jindex[ba];
    }
    if (l == 0 && live) {
        R.nn[i] = nn;
        R.flags[i] = refine ? 1 : 0;
    }
    if (done_word) {
        // Completion word: each workgroup releases its results to host memory
        // (system scope) before its ticket; the one that draws the last
        // ticket stores seq, which the host polls instead of an event.
        __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            if (atomicAdd(ticket, 1u) == gridDim.x - 1) {
                __hip_atomic_store(done_word, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                atomicExch(ticket, 0u);   // ready for the next launch
            }
        }
    }
}

// Lanes per i for the reduction: about two splits per lane, 4 to 64 lanes.
int reduce_lanes(int nsplit)
{
    int L = 4;
    while (L < 64 && 2 * L < nsplit) L *= 2;
    return L;
}

hipError_t launch_reduce(hipStream_t s, const Partials &p, int nsplit, int ni,
                         int nj, const IBlock &ib, float eps2,
                         int tile, const double *xp64, int cap, const int *jindex,
                         const Results &r, int lanes, const NearTile &nt,
                         const DoneWord *done)
{
    if (ni <= 0) return hipSuccess;
    if (nt.jchunk <= 0 || nt.jchunk % tile || nt.wpg < 1) return hipErrorInvalidValue;
    if (lanes <= 0) lanes = reduce_lanes(nsplit);
    const int per_wg = kReduceWaves * (64 / lanes);
    const dim3 grid((ni + per_wg - 1) / per_wg), block(64 * kReduceWaves);
    const DoneWord dw = done ? *done : DoneWord{nullptr, nullptr, 0};
    switch (lanes) {
    case 4:
        hipLaunchKernelGGL(k_reduce<2>, grid, block, 0, s, p, nsplit, ni, nj, ib, eps2,
                           tile, xp64, cap, jindex, r, nt, dw.ticket, dw.word, dw.seq);
        break;
    case 8:
        hipLaunchKernelGGL(k_reduce<3>, grid, block, 0, s, p, nsplit, ni, nj, ib, eps2,
                           tile, xp64, cap, jindex, r, nt, dw.ticket, dw.word, dw.seq);
        break;
    case 16:
        hipLaunchKernelGGL(k_reduce<4>, grid, block, 0, s, p, nsplit, ni, nj, ib, eps2,
                           tile, xp64, cap, jindex, r, nt, dw.ticket, dw.word, dw.seq);
        break;
    case 32:
        hipLaunchKernelGGL(k_reduce<5>, grid, block, 0, s, p, nsplit, ni, nj, ib, eps2,
                           tile, xp64, cap, jindex, r, nt, dw.ticket, dw.word, dw.seq);
        break;
    case 64:
        hipLaunchKernelGGL(k_reduce<6>, grid, block, 0, s, p, nsplit, ni, nj, ib, eps2,
                           tile, xp64, cap, jindex, r, nt, dw.ticket, dw.word, dw.seq);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Small ABI blocks in one launch (at most 12 i; kira's commonest blocks hold
// 1-7 nodes).  A block this small leaves most lanes idle in the
// i-per-lane force kernel and pays three launches (scatter, force,
// reduction).  Here each wave takes 64 j at a time, one per lane, predicts
// them (the predictor's arithmetic, so the same FP64 prediction) into a
// wave-private LDS area, and then splits its lanes into NIP groups of
// Q = 64 / NIP, one group per i: lane q of a group adds the pairs of j
// q, q + Q, ... of the 64 in FP64 (pair64, the pair law of
// hdyn_ev.C:2019-2036).  The sums carry no FP32 term and the nearest
// neighbour is the exact FP64 minimum (lowest address on ties,
// hdyn_inline.C:44).  A group combines its Q lanes once, at the end, by DPP
// row moves; the waves of a workgroup combine in wave order; each workgroup
// writes one partial per i, and the last workgroup to take a ticket adds
// them in workgroup order.  Every order is fixed by (nj, ni, grid), so the
// results are deterministic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void nn_min(double &bd, int &ba, double od, int oa)
{
    if (od < bd || (od == bd && oa < ba)) {
        bd = od;
        ba = oa;
    }
}

// v from another lane of the same row, by DPP control CTRL (all four rows
// and banks enabled; every pattern used reads a lane inside the row)
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_i32<CTRL>((int)(unsigned)u);
    const unsigned hi = (unsigned)dpp_i32<CTRL>((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int CTRL>
__device__ __forceinline__ void group_step_dpp(double F[7], double &bd, int &ba)
{
#pragma unroll
    for (int c = 0; c < 7; ++c) F[c] += dpp_f64<CTRL>(F[c]);
    const double od = dpp_f64<CTRL>(bd);
    const int oa = dpp_i32<CTRL>(ba);
    nn_min(bd, ba, od, oa);
}

__device__ __forceinline__ void group_step_xor(double F[7], double &bd, int &ba, int off)
{
#pragma unroll
    for (int c = 0; c < 7; ++c) F[c] += __shfl_xor(F[c], off);
    const double od = __shfl_xor(bd, off);
    const int oa = __shfl_xor(ba, off);
    nn_min(bd, ba, od, oa);
}

// Sum F and the nearest (bd, ba) over groups of Q consecutive lanes (Q a
// power of two, 4..64); afterwards every lane of a group holds the group's
// values (each step adds a lane and its partner in either order, which
// gives the same bits).  Pairs: xor 1 and xor 2 inside quads (quad_perm),
// then the other quad of 8 (row_half_mirror), the other half of the row of
// 16 (row_mirror), then across rows.
template <int Q>
__device__ __forceinline__ void group_reduce(double F[7], double &bd, int &ba)
{
    group_step_dpp<0xB1>(F, bd, ba);    // quad_perm [1,0,3,2]
    group_step_dpp<0x4E>(F, bd, ba);    // quad_perm [2,3,0,1]
    if (Q >= 8) group_step_dpp<0x141>(F, bd, ba);   // row_half_mirror
    if (Q >= 16) group_step_dpp<0x140>(F, bd, ba);  // row_mirror
    if (Q >= 32) group_step_xor(F, bd, ba, 16);
    if (Q >= 64) group_step_xor(F, bd, ba, 32);
}

struct JRawLite {   // the fields the predictor reads
    double x[3], v[3], a2[3], j6[3], k18[3], tj, m;
};

__device__ __forceinline__ JRawLite load_lite(const JRaw &r, int order)
{
    JRawLite o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.x[k] = r.x[k];
        o.v[k] = r.v[k];
        o.a2[k] = r.a2[k];
        o.j6[k] = r.j6[k];
        o.k18[k] = order >= 5 ? r.k18[k] : 0.0;
    }
    o.tj = r.tj;
    o.m = r.m;
    return o;
}

template <int NIP, int MD>
__global__ __launch_bounds__(kSmallThreads) void k_small(
    const SmallArgs<NIP, MD> B, JRaw *__restrict__ jraw, int *__restrict__ jindex,
    double *__restrict__ xp64, double *__restrict__ vp64, int cap, int chunks,
    double *__restrict__ part, unsigned *__restrict__ ticket, Results R,
    unsigned long long *done_flag, unsigned long long seq)
{
#pragma clang fp contract(off)
    constexpr int NW = kSmallThreads / 64;
    constexpr int Q = 64 / NIP;
    __shared__ double sj[NW][7][64];   // per wave: predicted x y z vx vy vz, m of its 64 j
    __shared__ double red[NW][NIP][kSmallComp];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int nj = B.nj, ni = B.ni;
    const int slot = lane / Q, q = lane - slot * Q;
    const int ic = slot < ni ? slot : ni - 1;
    const double xi = B.xi[ic][0], yi = B.xi[ic][1], zi = B.xi[ic][2];
    const double ui = B.vi[ic][0], vi = B.vi[ic][1], wi = B.vi[ic][2];
    const int ex = B.excl[ic];

    double F[7] = {0, 0, 0, 0, 0, 0, 0};
    double bd = __builtin_inf();
    int ba = INT_MAX;
    // chunk c of this wave: j in [jbase(c), jbase(c) + 64); the next chunk's
    // record is loaded while this one is summed
    auto jbase = [&](int c) { return ((c * G + (int)blockIdx.x) * NW + wave) * 64; };
    JRawLite nxt{};
    if (jbase(0) + lane < nj) nxt = load_lite(jraw[jbase(0) + lane], B.order);
    for (int c = 0; c < chunks; ++c) {
        const int j0 = jbase(c);
        if (j0 >= nj) break;   // uniform in the wave
        const int jn = min(64, nj - j0);
        const int j = j0 + lane;
        JRawLite cur = nxt;
        if (c + 1 < chunks && jbase(c + 1) + lane < nj)
            nxt = load_lite(jraw[jbase(c + 1) + lane], B.order);
        // pending updates inside this wave's 64 j: lane k tests update k, one
        // ballot names those in range (usually none), and for each the lane
        // that owns the address takes the record and stores it to HBM
        {
            const unsigned mine = lane < B.nd ? (unsigned)(B.daddr[lane] - j0) : ~0u;
            unsigned long long hits = __ballot(mine < (unsigned)jn);
            while (hits) {
                const int k = __builtin_ctzll(hits);
                hits &= hits - 1;
                if (lane == B.daddr[k] - j0) {
                    const JRaw &d = B.drec[k];
                    cur = load_lite(d, B.order);
                    jraw[j] = d;
                    jindex[j] = B.dindex[k];
                }
            }
        }
        if (lane < jn) {
            double xp[3], vp[3];
            predict_one(cur, B.ti, B.order, xp, vp);
            if (B.write_pred) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    xp64[(size_t)k * cap + j] = xp[k];
                    vp64[(size_t)k * cap + j] = vp[k];
                }
                vp64[3 * (size_t)cap + j] = cur.m;   // m64
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sj[wave][k][lane] = xp[k];
                sj[wave][3 + k][lane] = vp[k];
            }
            sj[wave][6][lane] = cur.m;
        }
        wave_lds_sync();
#pragma unroll 2
        for (int t = q; t < jn; t += Q) {
            const int jj = j0 + t;
            if (jj != ex) {
                const double dx = sj[wave][0][t] - xi;
                const double dy = sj[wave][1][t] - yi;
                const double dz = sj[wave][2][t] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz;
                pair64_r2(dx, dy, dz, sj[wave][3][t] - ui, sj[wave][4][t] - vi,
                          sj[wave][5][t] - wi, sj[wave][6][t], d2 + B.eps2, F);
                nn_min(bd, ba, d2, jj);
            }
        }
        wave_lds_sync();   // the next chunk overwrites the staging area
    }
    group_reduce<Q>(F, bd, ba);
    if (q == 0 && slot < ni) {
#pragma unroll
        for (int c = 0; c < 7; ++c) red[wave][slot][c] = F[c];
        red[wave][slot][7] = bd;
        red[wave][slot][8] = (double)ba;
    }
    __syncthreads();
    // wave 0 writes this workgroup's partial per i (waves added in order),
    // publishes it device-wide and takes a ticket; the last workgroup sees
    // every other's partials
    if (wave == 0) {
        for (int v = lane; v < ni * 8; v += 64) {
            const int i = v >> 3, c = v & 7;
            double *col = part + ((size_t)i * kSmallComp + c) * G + blockIdx.x;
            if (c < 7) {
                double s = red[0][i][c];
                for (int w = 1; w < NW; ++w) s += red[w][i][c];
                col[0] = s;
            } else {
                double d = red[0][i][7];
                int a = (int)red[0][i][8];
                for (int w = 1; w < NW; ++w) nn_min(d, a, red[w][i][7], (int)red[w][i][8]);
                col[0] = d;
                col[G] = (double)a;   // component 8
            }
        }
        // release: the partials are visible device-wide before the ticket
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) is_last = atomicAdd(ticket, 1u) == (unsigned)(G - 1);
    }
    __syncthreads();
    if (!is_last) return;
    // acquire: the last workgroup sees every partial released before its ticket
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // The last workgroup: NIP groups of QF threads, one group per i; thread
    // fq of a group adds workgroups fq, fq + QF, ... in order, then the group
    // combines its lanes (DPP) and, when it spans several waves, their sums
    // in wave order through LDS.
    constexpr int QF = kSmallThreads / NIP;
    const int fs = threadIdx.x / QF, fq = threadIdx.x - fs * QF;
    double S[7] = {0, 0, 0, 0, 0, 0, 0};
    double sd = __builtin_inf();
    int sa = INT_MAX;
    if (fs < ni) {
        // U workgroups' partials per round, all loads issued before the adds
        // (the partials come from other XCDs: every round is a trip to
        // memory); the adds keep workgroup order fq, fq + QF, ...  Groups of
        // 256 threads or more see at most two rounds and keep the registers.
        constexpr int U = QF >= 256 ? 1 : 4;
        const double *col = part + (size_t)fs * kSmallComp * G;
        for (int b0 = fq; b0 < G; b0 += U * QF) {
            double v[U][kSmallComp];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u * QF;
#pragma unroll
                for (int c = 0; c < kSmallComp; ++c)
                    v[u][c] = b < G ? col[(size_t)c * G + b] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (b0 + u * QF >= G) break;
#pragma unroll
                for (int c = 0; c < 7; ++c) S[c] += v[u][c];
                nn_min(sd, sa, v[u][7], (int)v[u][8]);
            }
        }
    }
    group_reduce<(QF < 64 ? QF : 64)>(S, sd, sa);
    if (QF > 64) {
        // red[][0] is free again: every wave's partials were read above
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 7; ++c) red[wave][0][c] = S[c];
            red[wave][0][7] = sd;
            red[wave][0][8] = (double)sa;
        }
        __syncthreads();
        if (fq == 0) {
#pragma unroll 1
            for (int w = 1; w < QF / 64; ++w) {
#pragma unroll
                for (int c = 0; c < 7; ++c) S[c] += red[wave + w][0][c];
                nn_min(sd, sa, red[wave + w][0][7], (int)red[wave + w][0][8]);
            }
        }
    }
    if (fq == 0 && fs < ni) {
        const int i = fs;
        int nn = 0;
        if (sa != INT_MAX) {
            nn = jindex[sa];
            // an index this launch wrote may not be visible here yet
            for (int k = 0; k < B.nd; ++k)
                if (B.daddr[k] == sa) nn = B.dindex[k];
        }
        R.acc[3 * i + 0] = S[0];
        R.acc[3 * i + 1] = S[1];
        R.acc[3 * i + 2] = S[2];
        R.jerk[3 * i + 0] = S[3];
        R.jerk[3 * i + 1] = S[4];
        R.jerk[3 * i + 2] = S[5];
        R.pot[i] = S[6];
        R.nn[i] = nn;
        R.flags[i] = 0;
    }
    if (done_flag) {
        // release at system scope: each writer's results reach host memory
        // before the completion word
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(done_flag, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (threadIdx.x == 0) atomicExch(ticket, 0u);   // ready for the next launch
}

// Workgroups and 64-j chunks per wave for nj j-particles: about two chunks per
// wave, at most kSmallMaxWG workgroups.  Fewer workgroups mean fewer partials
// for the last one to read and fewer tickets; measured per block (ni = 1 / 4 /
// 8): nj = 65536 19.2 / 22.4 / 27.7 us with 128 workgroups against 20.6 /
// 23.5 / 29.1 with 256; nj = 262144 26.9 / 32.8 / 38.7 us with 256 against
// 28.7 / 34.6 / 44.0 with 512 and 29.4 / 36.3 / 43.8 with 128.
void small_grid(int nj, int &wgs, int &chunks)
{
    wgs = std::min(kSmallMaxWG, std::max(1, (nj + 2 * kSmallThreads - 1) / (2 * kSmallThreads)));
    chunks = (nj + wgs * kSmallThreads - 1) / (wgs * kSmallThreads);
}

size_t small_part_doubles() { return (size_t)kSmallMaxI * kSmallComp * kSmallMaxWG; }

// One launch of k_small<N, MD>: the block copied into its argument form (N
// i-slots, room for MD >= b.nd pending records).
template <int N, int MD>
hipError_t launch_small_t(hipStream_t s, const SmallBlock &b, dim3 grid, JRaw *jraw,
                          int *jindex, double *xp64, double *vp64, int cap, int chunks,
                          double *part, unsigned *ticket, const Results &r,
                          unsigned long long *done_flag, unsigned long long seq)
{
    static_assert(N <= kSmallMaxI && MD <= kSmallMaxDirty, "argument form larger than the block");
    SmallArgs<N, MD> a;
    a.ti = b.ti;
    a.eps2 = b.eps2;
    a.ni = b.ni;
    a.nj = b.nj;
    a.nd = b.nd;
    a.order = b.order;
    a.write_pred = b.write_pred;
    a.pad = 0;
    for (int k = 0; k < N; ++k) {
        for (int c = 0; c < 3; ++c) {
            a.xi[k][c] = b.xi[k][c];
            a.vi[k][c] = b.vi[k][c];
        }
        a.excl[k] = b.excl[k];
    }
    for (int k = 0; k < b.nd; ++k) {
        a.daddr[k] = b.daddr[k];
        a.dindex[k] = b.dindex[k];
        a.drec[k] = b.drec[k];
    }
    hipLaunchKernelGGL((k_small<N, MD>), grid, dim3(kSmallThreads), 0, s, a, jraw, jindex, xp64,
                       vp64, cap, chunks, part, ticket, r, done_flag, seq);
    return hipGetLastError();
}

hipError_t launch_small(hipStream_t s, const SmallBlock &b, JRaw *jraw, int *jindex,
                        double *xp64, double *vp64, int cap, double *part, unsigned *ticket,
                        const Results &r, unsigned long long *done_flag, unsigned long long seq)
{
    if (b.ni < 1 || b.ni > kSmallMaxI || b.nj < 1 || b.nj > cap || b.nd < 0 ||
        b.nd > kSmallMaxDirty)
        return hipErrorInvalidValue;
    for (int k = 0; k < b.nd; ++k)
        if (b.daddr[k] < 0 || b.daddr[k] >= b.nj) return hipErrorInvalidValue;
    int wgs, chunks;
    small_grid(b.nj, wgs, chunks);
    const dim3 grid(wgs);
    int nip = 1;
    while (nip < b.ni) nip *= 2;
    const bool few = b.nd <= kSmallFewDirty;
#define G6_SMALL_LAUNCH(N)                                                                  \
    (few ? launch_small_t<N, kSmallFewDirty>(s, b, grid, jraw, jindex, xp64, vp64, cap,     \
                                             chunks, part, ticket, r, done_flag, seq)       \
         : launch_small_t<N, kSmallMaxDirty>(s, b, grid, jraw, jindex, xp64, vp64, cap,     \
                                             chunks, part, ticket, r, done_flag, seq))
    switch (nip) {
    case 1: return G6_SMALL_LAUNCH(1);
    case 2: return G6_SMALL_LAUNCH(2);
    case 4: return G6_SMALL_LAUNCH(4);
    case 8: return G6_SMALL_LAUNCH(8);
    case 16: return G6_SMALL_LAUNCH(16);
    default: return hipErrorInvalidValue;
    }
#undef G6_SMALL_LAUNCH
}

// ---------------------------------------------------------------------------
// Potential only (kira's energy path, pot_only in hdyn_grape6.C:477-551):
// the pair law's pot term, -m/r, with the mapping, double-single
// positions, self-exclusion ballot and FP32-tile / FP64-fold sums of
// k_force.  About 15 VALU instructions per pair instead of ~38.
// part[split][ld] receives one FP64 partial per i; k_pot_reduce sums the
// splits in order.
// ---------------------------------------------------------------------------
template <bool MASK>
__device__ __forceinline__ void pair_pot(const JPred &J, int jj, unsigned rel, float xih,
                                         float yih, float zih, float xil, float yil,
                                         float zil, float eps2, float &p)
{
#pragma clang fp contract(off)
    float dx, dy, dz;
    const float r2 = ds_r2(J.xh, J.yh, J.zh, J.xl, J.yl, J.zl, xih, yih, zih, xil, yil, zil,
                           eps2, dx, dy, dz);
    float rinv = __builtin_amdgcn_rsqf(r2);
    if (MASK) rinv = (unsigned)jj == rel ? 0.0f : rinv;
    p = __builtin_fmaf(-J.m, rinv, p);
}

template <bool MASK>
__device__ __forceinline__ float tile_pot(const JPred *__restrict__ jp, int tn, unsigned rel,
                                          float xih, float yih, float zih, float xil,
                                          float yil, float zil, float eps2)
{
    float p = 0.0f;
    int jj = 0;
    for (; jj + 4 <= tn; jj += 4) {
        const JPred *g = jp + jj;
        pair_pot<MASK>(g[0], jj + 0, rel, xih, yih, zih, xil, yil, zil, eps2, p);
        pair_pot<MASK>(g[1], jj + 1, rel, xih, yih, zih, xil, yil, zil, eps2, p);
        pair_pot<MASK>(g[2], jj + 2, rel, xih, yih, zih, xil, yil, zil, eps2, p);
        pair_pot<MASK>(g[3], jj + 3, rel, xih, yih, zih, xil, yil, zil, eps2, p);
    }
    for (; jj < tn; ++jj)
        pair_pot<MASK>(jp[jj], jj, rel, xih, yih, zih, xil, yil, zil, eps2, p);
    return p;
}

__global__ __launch_bounds__(64) void k_pot(
    const JPred *__restrict__ jpred, int nj, IBlock ib, int ni, float eps2,
    int jchunk, int tile, int n_igroups, double *__restrict__ part, int ld, int remap)
{
    int b = blockIdx.x;
    if (remap) {
        const int per = gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    const int split = b / n_igroups;
    const int ig = b - split * n_igroups;
    const int i = ig * 64 + threadIdx.x;
    const int ic = i < ni ? i : ni - 1;
    const float xih = ib.xh[ic], yih = ib.yh[ic], zih = ib.zh[ic];
    const float xil = ib.xl[ic], yil = ib.yl[ic], zil = ib.zl[ic];
    const int excl = i < ni ? ib.excl[ic] : -1;
    const float eps2v = to_vgpr(eps2);
    const int jbeg = split * jchunk;
    const int jend = min(jbeg + jchunk, nj);
    double P = 0.0;
    for (int t = jbeg; t < jend; t += tile) {
        const int tn = min(tile, jend - t);
        const unsigned rel = (unsigned)(excl - t);
        float p;
        if (__ballot(rel < (unsigned)tn) == 0)
            p = tile_pot<false>(jpred + t, tn, rel, xih, yih, zih, xil, yil, zil, eps2v);
        else
            p = tile_pot<true>(jpred + t, tn, rel, xih, yih, zih, xil, yil, zil, eps2v);
        P += (double)p;
    }
    if (i < ni) part[(size_t)split * ld + i] = P;
}

__global__ __launch_bounds__(256) void k_pot_reduce(const double *__restrict__ part, int nsplit,
                                                    int ld, int ni, double *__restrict__ pot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * ld + i];
    pot[i] = s;
}

hipError_t launch_pot(hipStream_t s, const JPred *jpred, int nj, const IBlock &ib, int ni,
                      float eps2, int nsplit, int jchunk, int tile, double *part, double *pot,
                      bool xcd_remap)
{
    if (ni <= 0) return hipSuccess;
    const int n_igroups = (ni + 63) / 64;
    const int nwg = n_igroups * nsplit;
    const int remap = (xcd_remap && nwg % 8 == 0) ? 1 : 0;
    hipLaunchKernelGGL(k_pot, dim3(nwg), dim3(64), 0, s, jpred, nj, ib, ni, eps2, jchunk,
                       tile, n_igroups, part, ib.ld, remap);
    hipLaunchKernelGGL(k_pot_reduce, dim3((ni + 255) / 256), dim3(256), 0, s, part, nsplit,
                       ib.ld, ni, pot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k-th nearest neighbours and local densities (compute_density,
// src/node/dyn/util/compute_density.C:69-226, the k = 12 density kira asks
// for in grape6_calculate_densities, hdyn_grape6.C:3604, 4283-4449).
//
// Distances are FP64, d^2 = (dx*dx + dy*dy) + dz*dz unfused, as Starlab's
// vec dot product (include/starlab_vector.h:68-70), so the neighbour sets
// and the k-th distance are exact.  Each lane keeps the k nearest j of its
// i (self excluded) as a list sorted by d^2, ties in ascending address
// (a later j is inserted after equal ones, as compute_density.C:165-176
// does).  j is wave-uniform (scalar loads); a j-split writes its partial
// list [split][q][ld]; k_knn_merge inserts the splits' lists in split order
// (= ascending address), then forms
//   density = (sum of the masses of the k-1 nearest) / (4/3 pi d_k^3)
// (compute_density.C:200-210; 0 when d_k = 0).
//
// The list lives in registers: KMAX slots, fully unrolled, with the runtime
// k <= KMAX only deciding which slot is the threshold.
// ---------------------------------------------------------------------------
template <int KMAX>
struct KList {
    double d[KMAX];
    int a[KMAX];
    double kth;
};

template <int KMAX>
__device__ __forceinline__ void klist_init(KList<KMAX> &L)
{
#pragma unroll
    for (int q = 0; q < KMAX; ++q) { L.d[q] = __builtin_inf(); L.a[q] = -1; }
    L.kth = __builtin_inf();
}

// Insert (d, a) if it is nearer than the k-th entry.  Stable: an entry
// goes after the ones with equal d already present.
template <int KMAX>
__device__ __forceinline__ void klist_insert(KList<KMAX> &L, int k, double d, int a)
{
    bool moved = false;
    double cd = d;
    int ca = a;
#pragma unroll
    for (int q = 0; q < KMAX; ++q) {
        const bool sw = moved || cd < L.d[q];
        const double td = L.d[q];
        const int ta = L.a[q];
        L.d[q] = sw ? cd : td;
        L.a[q] = sw ? ca : ta;
        cd = sw ? td : cd;
        ca = sw ? ta : ca;
        moved = sw;
    }
    double kth = L.d[0];
#pragma unroll
    for (int q = 1; q < KMAX; ++q) kth = (q == k - 1) ? L.d[q] : kth;
    L.kth = kth;
}

template <int KMAX>
__global__ __launch_bounds__(64) void k_knn(
    const double *__restrict__ xp64, int cap, int nj, int i_first, int ni, int k,
    int jchunk, int n_igroups, double *__restrict__ out_d, int *__restrict__ out_a, int ld)
{
#pragma clang fp contract(off)
    const int split = blockIdx.x / n_igroups;
    const int ig = blockIdx.x - split * n_igroups;
    const int i = ig * 64 + threadIdx.x;
    const int self = i_first + (i < ni ? i : ni - 1);
    const double xi = xp64[self], yi = xp64[(size_t)cap + self];
    const double zi = xp64[2 * (size_t)cap + self];
    const double *__restrict__ X = xp64;
    const double *__restrict__ Y = xp64 + cap;
    const double *__restrict__ Z = xp64 + 2 * (size_t)cap;
    KList<KMAX> L;
    klist_init(L);
    const int jbeg = split * jchunk;
    const int jend = min(jbeg + jchunk, nj);
    auto test = [&](int j, double xj, double yj, double zj) {
        const double dx = xj - xi;
        const double dy = yj - yi;
        const double dz = zj - zi;
        const double d2 = (dx * dx + dy * dy) + dz * dz;
        if (d2 < L.kth && j != self) klist_insert(L, k, d2, j);
    };
    int j = jbeg;
    // 8 j per step through wide scalar loads (one load latency per 8 j)
    for (; j + 8 <= jend; j += 8) {
        double XX[8], YY[8], ZZ[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            XX[u] = X[j + u];
            YY[u] = Y[j + u];
            ZZ[u] = Z[j + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) test(j + u, XX[u], YY[u], ZZ[u]);
    }
    for (; j < jend; ++j) test(j, X[j], Y[j], Z[j]);
    if (i < ni) {
#pragma unroll
        for (int q = 0; q < KMAX; ++q) {
            if (q < k) {
                const size_t o = ((size_t)split * KMAX + q) * ld + i;
                out_d[o] = L.d[q];
                out_a[o] = L.a[q];
            }
        }
    }
}

template <int KMAX>
__global__ __launch_bounds__(64) void k_knn_merge(
    const double *__restrict__ in_d, const int *__restrict__ in_a, int nsplit, int ld, int ni,
    int k, const JRaw *__restrict__ jraw, double *__restrict__ density,
    double *__restrict__ dk2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    KList<KMAX> L;
    klist_init(L);
    for (int s = 0; s < nsplit; ++s) {
        for (int q = 0; q < k; ++q) {
            const size_t o = ((size_t)s * KMAX + q) * ld + i;
            const double d = in_d[o];
            if (!(d < L.kth)) break;            // the split's list is sorted
            klist_insert(L, k, d, in_a[o]);
        }
    }
    double mass = 0.0;   // nearest first, as compute_density.C:207-208
#pragma unroll
    for (int q = 0; q < KMAX; ++q) {
        if (q < k - 1 && L.a[q] >= 0) mass += jraw[L.a[q]].m;
    }
    const double d2 = L.kth;
    double rho = 0.0;
    if (d2 > 0.0) {
        const double volume = (4.0 / 3.0) * 3.14159265358979323846 * pow(d2, 1.5);
        if (volume > 0.0) rho = mass / volume;
    }
    density[i] = rho;
    dk2[i] = d2;
}

int knn_kmax(int k) { return k <= 16 ? 16 : 32; }

hipError_t launch_knn(hipStream_t s, const double *xp64, int cap, int nj, int i_first, int ni,
                      int k, int nsplit, int jchunk, double *part_d, int *part_a, int ld,
                      const JRaw *jraw, double *density, double *dk2)
{
    if (ni <= 0) return hipSuccess;
    if (k < 2 || k > 32) return hipErrorInvalidValue;
    const int n_igroups = (ni + 63) / 64;
    const dim3 grid(n_igroups * nsplit), mgrid((ni + 63) / 64);
    if (knn_kmax(k) == 16) {
        hipLaunchKernelGGL(k_knn<16>, grid, dim3(64), 0, s, xp64, cap, nj, i_first, ni, k,
                           jchunk, n_igroups, part_d, part_a, ld);
        hipLaunchKernelGGL(k_knn_merge<16>, mgrid, dim3(64), 0, s, part_d, part_a, nsplit, ld,
                           ni, k, jraw, density, dk2);
    } else {
        hipLaunchKernelGGL(k_knn<32>, grid, dim3(64), 0, s, xp64, cap, nj, i_first, ni, k,
                           jchunk, n_igroups, part_d, part_a, ld);
        hipLaunchKernelGGL(k_knn_merge<32>, mgrid, dim3(64), 0, s, part_d, part_a, nsplit, ld,
                           ni, k, jraw, density, dk2);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Neighbour lists (FP64, exact).  The test is
//   ((e2[i] + dx*dx) + dy*dy) + dz*dz < h2[i]
// with e2 = 0 for the GRAPE-6 front end (0 + a = a exactly: the unsoftened
// d^2 < h2) and e2 = 2 eps2, h2 = 2 h2 for the GRAPE-4 front end, the
// operation order of the reference's IEEE GRAPE-4 simulator
// (harp3fastsim.c:289-303).  Lists are in ascending address; counts are the
// true counts, at most nb_cap entries per list are stored.
//
// Three passes over (64 listed i-slots x j-split) waves.  Lanes are slots and
// j is wave-uniform (scalar loads of the FP64 predicted positions), so each j
// is read once per 64 slots, and the j range is split over enough waves to
// fill the chip even for a single slot:
//   k_nl<false>        hits per (slot, split)                     -> cnt
//   k_nl_scan_splits   one wave per slot: exclusive offsets over its splits,
//                      the slot's true count and its clipped length
//   k_nl_scan_slots    one workgroup: packed bases of the clipped lists
//   k_nl<true>         the same sweep, writing each hit at its place in one
//                      packed array (slot s's list at base[s], nb_cap at most)
//
// FP32 filter.  The exact test costs ~12 FP64 instructions per (i, j); both
// sweeps first compare an FP32 r^2 against T = h2 - e2 and take the FP64
// test only inside a band around T that bounds the FP32 error, so the lists
// are the FP64 test's.  k_nl_prep writes FP32 copies of the j positions and,
// per 8 j, the largest |coordinate| m8.  With u = 2^-24, each FP32 difference
// is within delta = u (2 + u) (|x_j| + |x_i|) / (1 - u) <= 2^-22 (m8 + X_i) of
// the exact one, so |r| moves by at most sqrt(3) delta, and with the FP32
// rounding of the squares (3 u relative)
//   r2 >= T  =>  r2_32 >= T - (2 sqrt(3) sqrt(T) delta + 3 u T)
//   r2 <  T  =>  r2_32 <  T + (2 sqrt(3) sqrt(T) delta + 3 delta^2 + 3 u T)
// The band B = 4 sqrt(T) delta + 4 delta^2 + 2^-20 T + 2^-48 (h2 + e2)
// covers both, the FP32 rounding of T and B, and the FP64 test's own
// rounding (relative 2^-52 of h2 + e2).  r2_32 < T - B is a hit, r2_32 >
// T + B is not, and between them the FP64 test decides.
// ---------------------------------------------------------------------------
constexpr int kNlGroup = 8;   // j per step of the list sweeps, and per m8 entry

__global__ __launch_bounds__(256) void k_nl_prep(const double *__restrict__ xp64, int cap,
                                                 int nj, float *__restrict__ xf,
                                                 float *__restrict__ yf,
                                                 float *__restrict__ zf,
                                                 float *__restrict__ m8)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (j < nj) {
        const float a = (float)xp64[j];
        const float b = (float)xp64[(size_t)cap + j];
        const float c = (float)xp64[2 * (size_t)cap + j];
        xf[j] = a;
        yf[j] = b;
        zf[j] = c;
        m = fmaxf(fabsf(a), fmaxf(fabsf(b), fabsf(c)));
    }
    // groups of kNlGroup consecutive j are consecutive lanes of one wave
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    m = fmaxf(m, __shfl_xor(m, 4));
    if ((j & (kNlGroup - 1)) == 0 && j < nj) m8[j / kNlGroup] = m;
}

template <bool WRITE>
__global__ __launch_bounds__(64) void k_nl(
    const double *__restrict__ xp64, int cap, const int *__restrict__ jindex, int nj,
    IBlock ib, const double *__restrict__ h2, const double *__restrict__ e2,
    const int *__restrict__ slots, int nslots, int include_self, int jchunk, int nsplit,
    int *__restrict__ cnt, const int *__restrict__ base, int nb_cap, int *__restrict__ packed,
    const int *__restrict__ total, int host_cap, int *__restrict__ host_packed,
    const float *__restrict__ xf, const float *__restrict__ yf, const float *__restrict__ zf,
    const float *__restrict__ m8)
{
#pragma clang fp contract(off)
    const int n_sgroups = (nslots + 63) / 64;
    const int b = blockIdx.x;
    const int split = b / n_sgroups;
    const int sg = b - split * n_sgroups;
    const int s = sg * 64 + (int)threadIdx.x;
    const int sc = s < nslots ? s : nslots - 1;
    const int i = slots[sc];
    const double hh = h2[i], ee = e2[i];
    const double xi = ib.x64[i], yi = ib.y64[i], zi = ib.z64[i];
    const int excl = include_self ? -1 : ib.excl[i];
    const int jbeg = split * jchunk;
    const int jend = min(jbeg + jchunk, nj);
    const size_t c_at = (size_t)sc * nsplit + split;
    int pos = 0, lim = 0, out0 = 0;
    int *dst = packed;
    if (WRITE) {
        pos = cnt[c_at];   // offset of this split inside the slot's list
        lim = nb_cap;
        out0 = base[sc];
        // lists that fit the caller's pinned buffer go straight to the host
        if (*total <= host_cap) dst = host_packed;
    }
    const double *__restrict__ yp = xp64 + cap;
    const double *__restrict__ zp = xp64 + 2 * (size_t)cap;
    int c = 0;
    auto exact = [&](int j) {
        const double dx = xp64[j] - xi;
        const double dy = yp[j] - yi;
        const double dz = zp[j] - zi;
        const double d2 = ((ee + dx * dx) + dy * dy) + dz * dz;
        return d2 < hh && j != excl;
    };
    auto take = [&](int j, bool hit) {
        if (WRITE) {
            if (hit && s < nslots) {
                if (pos < lim) dst[out0 + pos] = jindex[j];
                ++pos;
            }
        } else {
            c += hit ? 1 : 0;
        }
    };
    // the lane's FP32 position and band terms (see above)
    const float xif = (float)xi, yif = (float)yi, zif = (float)zi;
    const float Xi = fmaxf(fabsf(xif), fmaxf(fabsf(yif), fabsf(zif)));
    const double Td = hh - ee;
    const float T = (float)Td;
    const float sT = Td > 0.0 ? sqrtf(T) : 0.0f;
    const float Bfix = T * 0x1p-20f + (float)(hh + ee) * 0x1p-48f;
    // d2 >= ee >= hh when T <= 0: no j can be a hit
    const bool none = !(Td > 0.0);
    int j = jbeg;
    // kNlGroup j per step (jbeg is a multiple of kNlGroup): their FP32
    // coordinates arrive as scalar loads issued together, one step ahead of
    // the step being tested (the last step reloads its own j)
    float X[kNlGroup], Y[kNlGroup], Z[kNlGroup], M = 0.0f;
    auto fetch = [&](int jf, float *Xs, float *Ys, float *Zs, float &Ms) {
#pragma unroll
        for (int u = 0; u < kNlGroup; ++u) {
            Xs[u] = xf[jf + u];
            Ys[u] = yf[jf + u];
            Zs[u] = zf[jf + u];
        }
        Ms = m8[jf / kNlGroup];
    };
    if (j + kNlGroup <= jend) fetch(j, X, Y, Z, M);
    for (; j + kNlGroup <= jend; j += kNlGroup) {
        float Xn[kNlGroup], Yn[kNlGroup], Zn[kNlGroup], Mn;
        fetch(j + 2 * kNlGroup <= jend ? j + kNlGroup : j, Xn, Yn, Zn, Mn);
        const float d = 0x1p-22f * (M + Xi);
        const float B = 4.0f * sT * d + 4.0f * d * d + Bfix;
        const float lo = T - B, hi = T + B;
        unsigned in = 0, amb = 0;
#pragma unroll
        for (int u = 0; u < kNlGroup; ++u) {
            const float dx = X[u] - xif;
            const float dy = Y[u] - yif;
            const float dz = Z[u] - zif;
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const bool a = !none && !(r2 < lo) && !(r2 > hi);
            in |= (!none && r2 < lo && j + u != excl) ? 1u << u : 0u;
            amb |= a ? 1u << u : 0u;
        }
        if (__ballot(amb != 0)) {
#pragma unroll
            for (int u = 0; u < kNlGroup; ++u)
                if ((amb >> u) & 1u) in |= exact(j + u) ? 1u << u : 0u;
        }
        if (WRITE) {
#pragma unroll
            for (int u = 0; u < kNlGroup; ++u) take(j + u, (in >> u) & 1u);
        } else {
            c += __popc(in);
        }
#pragma unroll
        for (int u = 0; u < kNlGroup; ++u) {
            X[u] = Xn[u];
            Y[u] = Yn[u];
            Z[u] = Zn[u];
        }
        M = Mn;
    }
    for (; j < jend; ++j) take(j, !none && exact(j));
    if (!WRITE && s < nslots) cnt[c_at] = c;
}

// In-place exclusive scan of a[0, n) by one workgroup of kScanThreads;
// returns the total (to every thread).
constexpr int kScanThreads = 1024;

__device__ int block_exclusive_scan(int *__restrict__ a, int n, int *part)
{
    const int t = threadIdx.x;
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int b0 = min(t * per, n), b1 = min(b0 + per, n);
    int local = 0;
    for (int k = b0; k < b1; ++k) local += a[k];
    // exclusive scan of the per-thread sums: inclusive scan inside each wave
    // (shuffles), then of the kScanThreads / 64 wave totals
    const int lane = t & 63, wave = t >> 6;
    int x = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) part[wave] = x;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) {
            const int v = part[w];
            part[w] = run;
            run += v;
        }
        part[kScanThreads] = run;
    }
    __syncthreads();
    int run = part[wave] + x - local;
    for (int k = b0; k < b1; ++k) {
        const int v = a[k];
        a[k] = run;
        run += v;
    }
    const int total = part[kScanThreads];
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(64) void k_nl_scan_splits(
    int *__restrict__ cnt, int nsplit, const int *__restrict__ slots, int nb_cap,
    int *__restrict__ counts, int *__restrict__ clip)
{
    // cnt is slot-major: [slot][split]
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    int *row = cnt + (size_t)s * nsplit;
    int carry = 0;
    for (int p0 = 0; p0 < nsplit; p0 += 64) {
        const int p = p0 + lane;
        const int v = p < nsplit ? row[p] : 0;
        int x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (p < nsplit) row[p] = carry + x - v;
        carry += __shfl(x, 63);
    }
    if (lane == 0) {
        counts[slots[s]] = carry;
        clip[s] = min(carry, nb_cap);
    }
}

// clip -> base (in place), total; both also to the host (host_info[nslots + 1])
__global__ __launch_bounds__(kScanThreads) void k_nl_scan_slots(int *__restrict__ clip,
                                                                int nslots,
                                                                int *__restrict__ total_out,
                                                                int *__restrict__ host_info)
{
    __shared__ int part[kScanThreads + 1];
    const int stored = block_exclusive_scan(clip, nslots, part);
    for (int s = threadIdx.x; s < nslots; s += kScanThreads) host_info[s] = clip[s];
    if (threadIdx.x == 0) {
        *total_out = stored;
        host_info[nslots] = stored;
    }
}

// j-split of the list passes: about 8192 waves in all (8 per SIMD: each wave
// walks its j one scalar-loaded step at a time), at least 16 j per wave,
// splits a multiple of kNlGroup j long
void nl_geometry(int nslots, int nj, int &nsplit, int &jchunk)
{
    const int n_sgroups = (nslots + 63) / 64;
    nsplit = std::max(1, std::min(8192 / n_sgroups, (nj + 15) / 16));
    jchunk = (nj + nsplit - 1) / nsplit;
    jchunk = (jchunk + kNlGroup - 1) / kNlGroup * kNlGroup;
    nsplit = (nj + jchunk - 1) / jchunk;
}

// scratch: cnt[nslots][nsplit], base[nslots], total, then (from a 16-B
// boundary) the FP32 j positions xf, yf, zf [nj] and m8[nj / kNlGroup]
static size_t nl_counts_ints(int nslots, int nsplit)
{
    return ((size_t)nslots * nsplit + nslots + 1 + 3) & ~(size_t)3;
}

size_t nl_scratch_ints(int nslots, int nj)
{
    int nsplit, jchunk;
    nl_geometry(nslots, nj, nsplit, jchunk);
    return nl_counts_ints(nslots, nsplit) + 3 * (size_t)nj + (nj + kNlGroup - 1) / kNlGroup;
}

hipError_t launch_nblist(hipStream_t s, const double *xp64, int cap, const int *jindex, int nj,
                         const IBlock &ib, const double *h2, const double *e2, const int *slots,
                         int nslots, int include_self, int nb_cap, int *counts, int *scratch,
                         int *packed, int *host_info, int *host_packed, int host_cap)
{
    if (nslots <= 0) return hipSuccess;
    int nsplit, jchunk;
    nl_geometry(nslots, nj, nsplit, jchunk);
    int *cnt = scratch;
    int *base = cnt + (size_t)nslots * nsplit;
    int *total = base + nslots;
    float *xf = reinterpret_cast<float *>(scratch + nl_counts_ints(nslots, nsplit));
    float *yf = xf + nj, *zf = yf + nj, *m8 = zf + nj;
    const dim3 grid(((nslots + 63) / 64) * nsplit);
    hipLaunchKernelGGL(k_nl_prep, dim3((nj + 255) / 256), dim3(256), 0, s, xp64, cap, nj, xf, yf,
                       zf, m8);
    hipLaunchKernelGGL(k_nl<false>, grid, dim3(64), 0, s, xp64, cap, jindex, nj, ib, h2, e2,
                       slots, nslots, include_self, jchunk, nsplit, cnt, base, nb_cap, packed,
                       total, host_cap, host_packed, xf, yf, zf, m8);
    hipLaunchKernelGGL(k_nl_scan_splits, dim3(nslots), dim3(64), 0, s, cnt, nsplit, slots,
                       nb_cap, counts, base);
    hipLaunchKernelGGL(k_nl_scan_slots, dim3(1), dim3(kScanThreads), 0, s, base, nslots, total,
                       host_info);
    hipLaunchKernelGGL(k_nl<true>, grid, dim3(64), 0, s, xp64, cap, jindex, nj, ib, h2, e2,
                       slots, nslots, include_self, jchunk, nsplit, cnt, base, nb_cap, packed,
                       total, host_cap, host_packed, xf, yf, zf, m8);
    return hipGetLastError();
}

}  // namespace g6
