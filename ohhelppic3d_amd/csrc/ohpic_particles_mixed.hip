// ohpic_particles_mixed.hip -- mixed-precision particle kernels for gfx950.
//
// Storage per binned particle (24 B): for each axis one 32-bit word holding
// the particle's cell relative to its tile origin in the top 8 bits (signed)
// and its in-cell offset as a 24-bit fixed-point fraction in the low 24 bits,
// and FP32 velocities.  The tile is the bin the particle sits in, so the cell
// index costs no storage; the offset f24 * 2^-24 converts exactly to FP32.
// Contiguous particle lists (host transfers, re-binning input, the overflow
// retry) carry the tile index in a seventh word.
//
// Gather, Boris, drift and the deposit weights run in FP32; the new offset is
// rounded to the 24-bit grid (carry into the cell) and the deposit uses that
// stored position; every deposit contribution is rounded into the same int64
// fixed-point charge grid the FP64 path uses, so rho, the Poisson solve, E
// and the diagnostics stay FP64.  The periodic boundary is an integer wrap of
// the cell index.  Work decomposition, LDS staging, deposit blocks and
// re-binning are those of the FP64 kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "ohpic_internal.hpp"
#include "ohpic_kernels.hpp"
#include "ohpic_particle_common.hpp"
#include "ohpic_pcg.hpp"

namespace ohpic {

namespace {

#ifndef OHPIC_MX_PREFETCH
#define OHPIC_MX_PREFETCH 1
#endif
// Minimum waves per SIMD of the fused mixed kernel: 5 (96 VGPRs) measured
// faster than the 4 the compiler picks on its own at 256^3 x 64 ppc.
#ifndef OHPIC_MX_MIN_WAVES
#define OHPIC_MX_MIN_WAVES 5
#endif
constexpr int kMxPrefetch = OHPIC_MX_PREFETCH;  // particles per lane loaded ahead

constexpr int kFracBits = 24;
constexpr uint32_t kFracOne = 1u << kFracBits;
constexpr uint32_t kFracMask = kFracOne - 1u;
constexpr float kFracScale = 16777216.0f;            // 2^24
constexpr float kFracUnit = 5.9604644775390625e-8f;  // 2^-24
constexpr float kMaxShift = 64.0f * kFracScale;      // largest drift per step: 64 cells

__device__ __forceinline__ uint32_t pos_word(int rel, uint32_t f24)
{
    return ((uint32_t)rel << kFracBits) | f24;
}
__device__ __forceinline__ int pos_rel(uint32_t w) { return (int32_t)w >> kFracBits; }
__device__ __forceinline__ float pos_frac(uint32_t w) { return (float)(w & kFracMask) * kFracUnit; }

struct PMx {
    uint32_t w[3];
    float v[3];
    int64_t id;
};

// binned form: buf[b][1..3] position words, buf[b][4..6] velocities
__device__ __forceinline__ PMx load_pmx(const SpecDev &S, int b, int64_t o, bool ids)
{
    PMx q;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        q.w[a] = gload((const uint32_t *)S.buf[b][1 + a], o);
        q.v[a] = gload((const float *)S.buf[b][4 + a], o);
    }
    q.id = ids ? gload(S.pid[b], o) : 0;
    return q;
}

__device__ __forceinline__ void store_pmx(const SpecDev &S, int b, int64_t o, const PMx &q,
                                          bool ids)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        gstore((uint32_t *)S.buf[b][1 + a], o, q.w[a]);
        gstore((float *)S.buf[b][4 + a], o, q.v[a]);
    }
    if (ids) gstore(S.pid[b], o, q.id);
}

// The bin of one tile in one buffer: array bases offset to the bin's first
// element (uniform per workgroup, so they live in scalar registers) and
// particles addressed by a 32-bit byte offset (a bin holds < 2^30 particles),
// which the global instructions take directly (no 64-bit address arithmetic
// per particle).
struct BinView {
    const char *w[3];
    const char *v[3];
    const int64_t *id;
};

__device__ __forceinline__ BinView bin_view(const SpecDev &S, int b, int64_t base, bool ids)
{
    BinView r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.w[a] = (const char *)((const uint32_t *)S.buf[b][1 + a] + base);
        r.v[a] = (const char *)((const float *)S.buf[b][4 + a] + base);
    }
    r.id = ids ? S.pid[b] + base : nullptr;
    return r;
}

__device__ __forceinline__ PMx load_bin(const BinView &bv, uint32_t slot, bool ids)
{
    const uint32_t off = slot * 4u;
    PMx q;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        q.w[a] = sload((const uint32_t *)(bv.w[a] + off), 0);
        q.v[a] = sload((const float *)(bv.v[a] + off), 0);
    }
    q.id = ids ? gload(bv.id, (int64_t)slot) : 0;
    return q;
}

__device__ __forceinline__ void store_bin(const BinView &bv, uint32_t slot, const PMx &q, bool ids)
{
    const uint32_t off = slot * 4u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sstore((uint32_t *)(bv.w[a] + off), 0, q.w[a]);
        sstore((float *)(bv.v[a] + off), 0, q.v[a]);
    }
    if (ids) gstore(const_cast<int64_t *>(bv.id), (int64_t)slot, q.id);
}

// list form: the binned words plus the tile index in buf[b][0]
__device__ __forceinline__ void store_list_mx(const SpecDev &S, int b, int64_t o, const PMx &q,
                                              int tile, bool ids)
{
    gstore((int32_t *)S.buf[b][0], o, (int32_t)tile);
    store_pmx(S, b, o, q, ids);
}

template <int WO>
__device__ __forceinline__ void cic_wf(float r, float &w0, float &w1)
{
    if (WO == 0) {
        w0 = r;
        w1 = 1.0f - r;
    } else {
        w0 = 1.0f - r;
        w1 = r;
    }
}

// g mod n: one add or subtract for a cell within a box length of [0, n)
// (every step of a particle slower than a box length per step), a division
// only beyond that
__device__ __forceinline__ int wrap_cell(int g, int n)
{
    if (g < 0 || g >= n) {
        g += g < 0 ? n : -n;
        if (g < 0 || g >= n) g = ((g % n) + n) % n;
    }
    return g;
}

__device__ __forceinline__ void route_back_mx(const SpecDev &S, int nbuf, int t, const PMx &q,
                                              bool ids, unsigned long long *flags)
{
    const unsigned long long old = gatomic_add(&S.count[nbuf][t], 1ull << 32);
    const int f = bin_front(old), b = bin_back(old);
    const int cap = gload(S.bin_cap, t);
    if (b + f < cap)
        store_pmx(S, nbuf, gload(S.bin_base, t) + (cap - 1 - b), q, ids);
    else
        gatomic_add(&flags[kFlagOverflow], 1ull);
}

struct LeaveBufMx {
    uint32_t w[3][kLeaveCap];
    float v[3][kLeaveCap];
    int64_t id[kLeaveCap];
    int32_t dest[kLeaveCap];
};

__device__ __forceinline__ void flush_leavers_mx(LeaveBufMx &lb, int cnt, const SpecDev &S,
                                                 int nbuf, bool ids, unsigned long long *flags)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int l = lane_id();
    if (l < cnt) {
        PMx q;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            q.w[a] = lb.w[a][l];
            q.v[a] = lb.v[a][l];
        }
        q.id = lb.id[l];
        route_back_mx(S, nbuf, lb.dest[l], q, ids, flags);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TX, int TY, int TZ, int WO, int MODE_>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(OHPIC_MX_MIN_WAVES, 8)))
    k_particles_mixed(const InstDev *__restrict__ insts, const WorkItem *__restrict__ items,
                      const int *__restrict__ parity, unsigned int *__restrict__ queue,
                      int nitems)
{
    // kModeFusedQueue: the fused step over a resident grid, each workgroup
    // claiming work items in list order from `queue` until none is left
    constexpr bool QUEUE = (MODE_ == kModeFusedQueue);
    constexpr int MODE = QUEUE ? (int)kModeFused : MODE_;
    constexpr bool LIST = (MODE == kModeFusedList || MODE == kModeBoundaryList);
    constexpr bool PUSH = (MODE == kModeFused || MODE == kModeGatherPush || MODE == kModeFusedList);
    constexpr bool REBIN = (MODE == kModeFused || MODE == kModeBoundary || LIST);
    constexpr bool DEP = (MODE == kModeFused || MODE == kModeDeposit || MODE == kModeFusedList);
    constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1, NE = EX * EY * EZ;
    constexpr int RX = TX + 3, RY = TY + 3, RZ = TZ + 3, NR = RX * RY * RZ;
    // E tile: (E_x, E_y) of a node side by side and E_z apart, so that the
    // gather reads the two x-neighbours of each (y, z) pair with one 2x64-bit
    // and one 2x32-bit LDS load at constant offsets from the particle's base
    // node (measured 3% faster per step than 16-byte nodes, whose loads share
    // fewer LDS banks, and than three separate arrays)
    __shared__ float2 sExy[PUSH ? NE : 1];
    __shared__ float sEz[PUSH ? NE : 1];
    __shared__ unsigned long long sR[DEP ? NR : 1];
    __shared__ LeaveBufMx sLeave[(REBIN && !LIST) ? kWaves : 1];
    __shared__ int sLcnt[kWaves];
    __shared__ double sRed[kWaves];
    __shared__ int sFront, sFrontLimit;
    __shared__ int64_t sListBase;

    __shared__ int sItem;
    for (int iter = 0;; ++iter) {
        int item;
        if (QUEUE) {
            if (threadIdx.x == 0) sItem = (int)atomicAdd(queue, 1u);
            __syncthreads();
            item = sItem;
            if (item >= nitems) break;
        } else {
            if (iter > 0) break;
            item = blockIdx.x;
        }
        const WorkItem it = items[item];
        const InstDev *__restrict__ I = insts + it.inst;
        const int par = parity[it.inst];
        const int nx = I->n[0], ny = I->n[1], nz = I->n[2];
        const int ntx = I->nt[0], nty = I->nt[1];
        const int tile = it.tile;
        const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
        const int lox = tx * TX, loy = ty * TY, loz = tz * TZ;
        const int tid = threadIdx.x;
        const int wv = tid >> 6;
        const bool ids = I->track_ids != 0;
        const bool excl = it.nchunks == 1;
        unsigned long long *__restrict__ flags = I->flags;

        if (PUSH) {
            const double *__restrict__ ex = I->e[0];
            const double *__restrict__ ey = I->e[1];
            const double *__restrict__ ez = I->e[2];
            for (int l = tid; l < NE; l += kBlock) {
                const int lx = l % EX, ly = (l / EX) % EY, lz = l / (EX * EY);
                const int gx = wrap_near(lox + lx, nx), gy = wrap_near(loy + ly, ny),
                          gz = wrap_near(loz + lz, nz);
                const int64_t g = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
                sExy[l] = make_float2((float)gload(ex, g), (float)gload(ey, g));
                sEz[l] = (float)gload(ez, g);
            }
        }
        if (DEP)
            for (int l = tid; l < NR; l += kBlock) sR[l] = 0ull;
        if (tid < kWaves) sLcnt[tid] = 0;
        __syncthreads();

        const float dt = (float)I->dt;
        const float dt2 = 0.5f * dt;
        const float dts = dt * kFracScale;  // cells per unit velocity, in 2^-24 units
        const float bx = (float)I->b0[0], by = (float)I->b0[1], bz = (float)I->b0[2];
        const int nbuf = par ^ 1;
        const unsigned long long lanemask_lt = (1ull << lane_id()) - 1ull;
        unsigned long long *__restrict__ racc =
            I->rho_acc[(MODE == kModeFused || MODE == kModeFusedList) ? nbuf : par];

        for (int s = 0; s < I->nspecies; ++s) {
            const SpecDev &S = I->sp[s];
            const int cap = gload(S.bin_cap, tile);
            const unsigned long long word = gload(S.count[par], tile);
            const int F = min(bin_front(word), cap);
            const int B = min(bin_back(word), cap - F);
            const int cnt = F + B;
            const int64_t base = gload(S.bin_base, tile);
            int per = (cnt + it.nchunks - 1) / it.nchunks;
            per = (per + 63) & ~63;
            const int start = it.chunk * per;
            const int end = min(cnt, start + per);
            if (REBIN) {
                // the previous species' particles are all stored before its
                // list base and front counters are replaced
                __syncthreads();
                if (tid == 0) {
                    int lim = cap;
                    if (LIST && end > start)  // one reservation per item and species
                        sListBase = (int64_t)gatomic_add(&S.count[nbuf][0],
                                                         (unsigned long long)(end - start));
                    if (!LIST && excl && cnt > 0) {
                        const unsigned long long old =
                            gatomic_add(&S.count[nbuf][tile], (unsigned long long)cnt);
                        lim = cap - bin_back(old);
                    }
                    sFront = 0;
                    sFrontLimit = lim;
                }
                __syncthreads();
            }
            const float qm = (float)S.qm;
            const double qfix = S.qfix;
            const float t0 = qm * bx * dt2, t1 = qm * by * dt2, t2 = qm * bz * dt2;
            const float tt = t0 * t0 + t1 * t1 + t2 * t2;
            const float s0 = 2.0f * t0 / (1.0f + tt), s1 = 2.0f * t1 / (1.0f + tt),
                        s2 = 2.0f * t2 / (1.0f + tt);
            double keo = 0.0, ken = 0.0;

            // the next kMxPrefetch particles of this lane are loaded ahead, so each
            // lane keeps that many particle reads in flight
            const BinView cur = bin_view(S, par, base, ids);
            const BinView nxt = bin_view(S, nbuf, base, ids);
            PMx ahead[kMxPrefetch];
#pragma unroll
            for (int k = 0; k < kMxPrefetch; ++k) {
                ahead[k] = PMx{{0u, 0u, 0u}, {0.0f, 0.0f, 0.0f}, 0};
                if (start + tid + k * kBlock < end)
                    ahead[k] = load_bin(cur, (uint32_t)bin_slot(start + tid + k * kBlock, F, B, cap),
                                        ids);
            }
            for (int p0 = start; p0 < end; p0 += kBlock) {
                const int p = p0 + tid;
                const bool active = p < end;
                PMx q = ahead[0];
#pragma unroll
                for (int k = 0; k + 1 < kMxPrefetch; ++k) ahead[k] = ahead[k + 1];
                if (p + kMxPrefetch * kBlock < end)
                    ahead[kMxPrefetch - 1] =
                        load_bin(cur, (uint32_t)bin_slot(p + kMxPrefetch * kBlock, F, B, cap), ids);
                // cell relative to this tile's origin and in-cell offset per axis
                int rx = pos_rel(q.w[0]), ry = pos_rel(q.w[1]), rz = pos_rel(q.w[2]);
                float ox = pos_frac(q.w[0]), oy = pos_frac(q.w[1]), oz = pos_frac(q.w[2]);
                if (PUSH && active) {
                    float wx0, wx1, wy0, wy1, wz0, wz1;
                    cic_wf<WO>(ox, wx0, wx1);
                    cic_wf<WO>(oy, wy0, wy1);
                    cic_wf<WO>(oz, wz0, wz1);
                    const int lx = min(max(rx, 0), TX - 1);
                    const int ly = min(max(ry, 0), TY - 1);
                    const int lz = min(max(rz, 0), TZ - 1);
                    // base node with 24-bit multiplies (full rate)
                    const unsigned nb =
                        __umul24((unsigned)lz, EX * EY) + __umul24((unsigned)ly, EX) + (unsigned)lx;
                    const float2 *__restrict__ exy = sExy + nb;
                    const float *__restrict__ ez0 = sEz + nb;
                    float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
                    const float wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int k = 0; k < 2; ++k) {
                                const float2 exy_n = exy[i + EX * (j + EY * k)];
                                const float w = wx[i] * wy[j] * wz[k];
                                e0 += exy_n.x * w;
                                e1 += exy_n.y * w;
                                e2 += ez0[i + EX * (j + EY * k)] * w;
                            }
                    const float vx = q.v[0], vy = q.v[1], vz = q.v[2];
                    keo += (double)(vx * vx + vy * vy + vz * vz);
                    const float um0 = vx + qm * e0 * dt2;
                    const float um1 = vy + qm * e1 * dt2;
                    const float um2 = vz + qm * e2 * dt2;
                    const float ua0 = um0 + (um1 * t2 - um2 * t1);
                    const float ua1 = um1 + (um2 * t0 - um0 * t2);
                    const float ua2 = um2 + (um0 * t1 - um1 * t0);
                    const float up0 = um0 + (ua1 * s2 - ua2 * s1);
                    const float up1 = um1 + (ua2 * s0 - ua0 * s2);
                    const float up2 = um2 + (ua0 * s1 - ua1 * s0);
                    q.v[0] = up0 + qm * e0 * dt2;
                    q.v[1] = up1 + qm * e1 * dt2;
                    q.v[2] = up2 + qm * e2 * dt2;
                    ken += (double)(q.v[0] * q.v[0] + q.v[1] * q.v[1] + q.v[2] * q.v[2]);
                    // drift in fixed point: the tile-relative position word (8.24)
                    // plus the displacement rounded to 2^-24 of a cell, exact in
                    // int32 while |v dt| < 64 cells (farther is reported)
                    const float sx = q.v[0] * dts, sy = q.v[1] * dts, sz = q.v[2] * dts;
                    const bool ok =
                        fabsf(sx) < kMaxShift && fabsf(sy) < kMaxShift && fabsf(sz) < kMaxShift;
                    if (!ok) {
                        gatomic_add(&flags[kFlagRange], 1ull);
                    } else {
                        q.w[0] = (uint32_t)((int32_t)q.w[0] + (int32_t)rintf(sx));
                        q.w[1] = (uint32_t)((int32_t)q.w[1] + (int32_t)rintf(sy));
                        q.w[2] = (uint32_t)((int32_t)q.w[2] + (int32_t)rintf(sz));
                    }
                    rx = pos_rel(q.w[0]);
                    ry = pos_rel(q.w[1]);
                    rz = pos_rel(q.w[2]);
                    ox = pos_frac(q.w[0]);
                    oy = pos_frac(q.w[1]);
                    oz = pos_frac(q.w[2]);
                    if (MODE == kModeGatherPush) {
                        // not re-binned: the cell stays relative to this tile
                        store_bin(cur, (uint32_t)bin_slot(p, F, B, cap), q, ids);
                    }
                }
                if (MODE == kModeGatherPush) continue;
                // global cell, periodic (in [0, n)); new tile and the cell
                // relative to it
                const int gx = wrap_cell(lox + rx, nx), gy = wrap_cell(loy + ry, ny),
                          gz = wrap_cell(loz + rz, nz);
                const int ctx_t = (int)((unsigned)gx / TX), cty = (int)((unsigned)gy / TY),
                          ctz = (int)((unsigned)gz / TZ);
                // tile counts are below 2^24 (axes <= 1024 cells, tiles >= 4)
                const int ntile = ctx_t + (int)__umul24((unsigned)cty, (unsigned)ntx) +
                                  (int)__umul24((unsigned)ctz, (unsigned)(ntx * nty));
                if (REBIN) {
                    q.w[0] = pos_word(gx - ctx_t * TX, q.w[0] & kFracMask);
                    q.w[1] = pos_word(gy - cty * TY, q.w[1] & kFracMask);
                    q.w[2] = pos_word(gz - ctz * TZ, q.w[2] & kFracMask);
                }
                if (DEP && active) {
                    float wx0, wx1, wy0, wy1, wz0, wz1;
                    cic_wf<WO>(ox, wx0, wx1);
                    cic_wf<WO>(oy, wy0, wy1);
                    cic_wf<WO>(oz, wz0, wz1);
                    const float wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
                    // node (x, y, z) of the deposit block is cell lo - 1 + (x, y, z)
                    // taken periodically, so the particle's cell relative to the
                    // tile plus one indexes it, wrapped or not
                    const unsigned dx = (unsigned)(rx + 1), dy = (unsigned)(ry + 1),
                                   dz = (unsigned)(rz + 1);
                    if (dx <= TX + 1 && dy <= TY + 1 && dz <= TZ + 1) {
                        unsigned long long *__restrict__ r0 = (unsigned long long *)(
                            (char *)sR + (__umul24(dz, RX * RY * 8u) + __umul24(dy, RX * 8u) +
                                          dx * 8u));
                        // the fixed-point charge of node (i, j, k) is
                        // round(qfix * (wx wy) * wz) with wx wy rounded to FP32 and
                        // the rest exact in FP64 (FP32 products of 24-bit
                        // mantissas, one rounding in the FMA)
                        const double wzd[2] = {(double)wz0, (double)wz1};
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 2; ++j) {
                                const double qxy = (double)(wx[i] * wy[j]) * qfix;
#pragma unroll
                                for (int k = 0; k < 2; ++k)
                                    atomicAdd(&r0[i + RX * (j + RY * k)],
                                              (unsigned long long)fix_round(wzd[k], qxy));
                            }
                    } else {
                        gatomic_add(&flags[kFlagFast], 1ull);
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 2; ++j) {
                                const double qxy = (double)(wx[i] * wy[j]) * qfix;
#pragma unroll
                                for (int k = 0; k < 2; ++k) {
                                    const int hx = (gx + i) % nx, hy = (gy + j) % ny,
                                              hz = (gz + k) % nz;
                                    const long long v = fix_round((double)wz[k], qxy);
                                    gatomic_add(&racc[hx + (int64_t)nx * (hy + (int64_t)ny * hz)],
                                                (unsigned long long)v);
                                }
                            }
                    }
                }
                if (LIST) {
                    // retry path: every particle to the species' contiguous
                    // list, in bin order inside this item's reserved range
                    if (active) store_list_mx(S, nbuf, sListBase + (p - start), q, ntile, ids);
                } else if (REBIN) {
                    const bool stay = active && ntile == tile;
                    const unsigned long long m = __ballot(stay);
                    if (m) {
                        const int leader = __ffsll((unsigned long long)m) - 1;
                        const int nst = (int)__popcll(m);
                        int b = 0, lim = 0;
                        if (excl) {
                            if (lane_id() == leader) b = atomicAdd(&sFront, nst);
                            lim = sFrontLimit;
                        } else if (lane_id() == leader) {
                            const unsigned long long old =
                                gatomic_add(&S.count[nbuf][tile], (unsigned long long)nst);
                            b = bin_front(old);
                            lim = cap - bin_back(old);
                        }
                        b = lane_bcast(b, leader);
                        lim = lane_bcast(lim, leader);
                        if (stay) {
                            const int slot = b + (int)__popcll(m & lanemask_lt);
                            if (slot < lim)
                                store_bin(nxt, (uint32_t)slot, q, ids);
                            else
                                gatomic_add(&flags[kFlagOverflow], 1ull);
                        }
                    }
                    const bool leave = active && ntile != tile;
                    const unsigned long long lm = __ballot(leave);
                    if (lm) {
                        const int nl = (int)__popcll(lm);
                        int c = sLcnt[wv];
                        if (c + nl > kLeaveCap) {
                            flush_leavers_mx(sLeave[wv], c, S, nbuf, ids, flags);
                            c = 0;
                        }
                        if (nl <= kLeaveCap) {
                            if (leave) {
                                const int slot = c + (int)__popcll(lm & lanemask_lt);
                                LeaveBufMx &lb = sLeave[wv];
#pragma unroll
                                for (int a = 0; a < 3; ++a) {
                                    lb.w[a][slot] = q.w[a];
                                    lb.v[a][slot] = q.v[a];
                                }
                                lb.id[slot] = q.id;
                                lb.dest[slot] = ntile;
                            }
                            c += nl;
                        } else if (leave) {
                            route_back_mx(S, nbuf, ntile, q, ids, flags);
                        }
                        if (lane_id() == 0) sLcnt[wv] = c;
                    }
                }
            }
            if (REBIN && !LIST) {
                flush_leavers_mx(sLeave[wv], sLcnt[wv], S, nbuf, ids, flags);
                __syncthreads();
                if (tid < kWaves) sLcnt[tid] = 0;
                if (tid == 0 && excl && cnt > sFront)
                    gatomic_add(&S.count[nbuf][tile],
                              (unsigned long long)(-(long long)(cnt - sFront)));
                __syncthreads();
            }
            if (PUSH) {
                const double so = block_sum(keo, sRed);
                const double sn = block_sum(ken, sRed);
                if (tid == 0) {
                    gstore(I->ke_partial, (int64_t)s * I->nitems + it.slot, 0.5 * S.mass * so);
                    gstore(I->ke_partial, (int64_t)(kMaxSpecies + s) * I->nitems + it.slot,
                           0.5 * S.mass * sn);
                }
            }
        }

        if (DEP) {
            __syncthreads();
            for (int l = tid; l < NR; l += kBlock) {
                const unsigned long long v = sR[l];
                if (v == 0ull) continue;
                const int lx = l % RX, ly = (l / RX) % RY, lz = l / (RX * RY);
                const int gx = wrap_near(lox - 1 + lx, nx);
                const int gy = wrap_near(loy - 1 + ly, ny);
                const int gz = wrap_near(loz - 1 + lz, nz);
                gatomic_add(&racc[gx + (int64_t)nx * (gy + (int64_t)ny * gz)], v);
            }
        }
        // the next item reuses the LDS blocks and sItem
        __syncthreads();
    }
}

// ---------------------------------------------------------------- binning
// A list entry (tile, words) names a cell inside the box: tile in range, the
// relative cell inside the tile, the global cell inside the (possibly partly
// covered) last tile, the fraction below 1 by construction.
__device__ __forceinline__ bool list_valid(const InstDev *I, int tile, const PMx &q, int tx_,
                                           int ty_, int tz_)
{
    if (tile < 0 || tile >= I->ntiles) return false;
    const int t3[3] = {tile % I->nt[0], (tile / I->nt[0]) % I->nt[1], tile / (I->nt[0] * I->nt[1])};
    const int T3[3] = {tx_, ty_, tz_};
    for (int a = 0; a < 3; ++a) {
        const int r = pos_rel(q.w[a]);
        if (r < 0 || r >= T3[a] || t3[a] * T3[a] + r >= I->n[a]) return false;
    }
    return true;
}

__global__ void k_bin_count_mx(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                               int64_t n, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    PMx q{{0u, 0u, 0u}, {0.0f, 0.0f, 0.0f}, 0};
    int t = 0;
    if (live) {
        q = load_pmx(S, lbuf, i, false);
        t = gload((const int32_t *)S.buf[lbuf][0], i);
    }
    const bool ok = live && list_valid(I, t, q, tx_, ty_, tz_);
    if (live && !ok) gatomic_add(&I->flags[kFlagRange], 1ull);
    (void)wave_counter_add(S.count[lbuf ^ 1], ok ? t : 0, ok);
}

__global__ void k_bin_scatter_mx(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                                 int64_t n, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const bool ids = I->track_ids != 0;
    PMx q{{0u, 0u, 0u}, {0.0f, 0.0f, 0.0f}, 0};
    int t = 0;
    if (live) {
        q = load_pmx(S, lbuf, i, ids);
        t = gload((const int32_t *)S.buf[lbuf][0], i);
    }
    const bool ok = live && list_valid(I, t, q, tx_, ty_, tz_);
    const unsigned long long old = wave_counter_add(S.count[lbuf ^ 1], ok ? t : 0, ok);
    if (!ok) return;
    const int slot = bin_front(old);
    if (slot < gload(S.bin_cap, t))
        store_pmx(S, lbuf ^ 1, gload(S.bin_base, t) + slot, q, ids);
    else
        gatomic_add(&I->flags[kFlagOverflow], 1ull);
}

__global__ void k_bins_to_list_mx(const InstDev *__restrict__ insts, int inst, int s, int src,
                                  int dst, const int *__restrict__ tiles, int t0,
                  const int64_t *__restrict__ list_off)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int t = tiles ? tiles[blockIdx.x] : t0 + (int)blockIdx.x;
    const unsigned long long w = S.count[src][t];
    const int cap = S.bin_cap[t];
    const int F = min(bin_front(w), cap);
    const int B = min(bin_back(w), cap - F);
    const int64_t ib = S.bin_base[t];
    const int64_t ob = list_off[blockIdx.x];
    const bool ids = I->track_ids != 0;
    for (int v = threadIdx.x; v < F + B; v += blockDim.x) {
        const PMx q = load_pmx(S, src, ib + bin_slot(v, F, B, cap), ids);
        store_list_mx(S, dst, ob + v, q, t, ids);
    }
}

__global__ void k_cell_counts_mx(const InstDev *__restrict__ insts, int inst, int s, int par,
                                 int32_t *__restrict__ counts, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int t = blockIdx.x;
    const unsigned long long w = S.count[par][t];
    const int cap = S.bin_cap[t];
    const int F = min(bin_front(w), cap);
    const int B = min(bin_back(w), cap - F);
    const int64_t b = S.bin_base[t];
    const int nx = I->n[0], ny = I->n[1];
    const int ox = (t % I->nt[0]) * tx_, oy = ((t / I->nt[0]) % I->nt[1]) * ty_,
              oz = (t / (I->nt[0] * I->nt[1])) * tz_;
    for (int v = threadIdx.x; v < F + B; v += blockDim.x) {
        const int64_t o = b + bin_slot(v, F, B, cap);
        const int cx = ox + pos_rel(((const uint32_t *)S.buf[par][1])[o]);
        const int cy = oy + pos_rel(((const uint32_t *)S.buf[par][2])[o]);
        const int cz = oz + pos_rel(((const uint32_t *)S.buf[par][3])[o]);
        atomicAdd(&counts[cx + (int64_t)nx * (cy + (int64_t)ny * cz)], 1);
    }
}

// ---------------------------------------------------------------- loaders
// FP64 position -> list entry.  The cell is the FP64 int() truncation; an
// offset that rounds up to 1 on the 24-bit grid is clamped below it, so the
// particle stays on the same side of the cell face as in FP64 (the mirrored
// CIC weights jump there).
__device__ __forceinline__ void store_loaded_mx(const InstDev *I, const SpecDev &S, int b,
                                                int64_t i, double x, double y, double z,
                                                double vx, double vy, double vz, int64_t id,
                                                bool ids, int tx_, int ty_, int tz_)
{
    const double p[3] = {x, y, z};
    const int T3[3] = {tx_, ty_, tz_};
    int tc[3];
    PMx q;
    for (int a = 0; a < 3; ++a) {
        const int c = (int)p[a];
        double f = rint((p[a] - (double)c) * 16777216.0);
        if (f > 16777215.0) f = 16777215.0;
        tc[a] = c / T3[a];
        q.w[a] = pos_word(c - tc[a] * T3[a], (uint32_t)f);
    }
    q.v[0] = (float)vx;
    q.v[1] = (float)vy;
    q.v[2] = (float)vz;
    q.id = id;
    store_list_mx(S, b, i, q, tc[0] + I->nt[0] * (tc[1] + I->nt[1] * tc[2]), ids);
}

__global__ void k_load_uniform_mx(const InstDev *__restrict__ insts, int inst, int s, int b,
                                  Pcg32Dev rng0, uint64_t offset, int64_t np, double vth,
                                  int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Pcg32Dev r = rng0;
    pcg_advance(r, offset + 7u * (uint64_t)i);
    const double x = pcg_double(r) * (double)I->n[0] + 0.0;
    const double y = pcg_double(r) * (double)I->n[1] + 0.0;
    const double z = pcg_double(r) * (double)I->n[2] + 0.0;
    double vx, vy, vz;
    maxwell3(r, vth, vx, vy, vz);
    store_loaded_mx(I, S, b, i, x, y, z, vx, vy, vz, i, I->track_ids != 0, tx_, ty_, tz_);
}

__global__ void k_load_hump_mx(const InstDev *__restrict__ insts, int inst, int s, int b,
                               Pcg32Dev rng0, uint64_t offset, uint64_t budget, int64_t np,
                               double amp, double sigma, double vth,
                               unsigned long long *__restrict__ fail, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int nx = I->n[0], ny = I->n[1], nz = I->n[2];
    Pcg32Dev r = rng0;
    pcg_advance(r, offset + (uint64_t)i * budget);
    const double cx = 0.5 * nx, cy = 0.5 * ny, cz = 0.5 * nz;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    const uint64_t maxc = (budget - 4u) / 4u;
    double x = cx, y = cy, z = cz;
    bool ok = false;
    for (uint64_t c = 0; c < maxc; ++c) {
        const double px = pcg_double(r) * (double)nx;
        const double py = pcg_double(r) * (double)ny;
        const double pz = pcg_double(r) * (double)nz;
        const double u = pcg_double(r);
        const double dx = px - cx, dy = py - cy, dz = pz - cz;
        const double f = (1.0 + amp * exp(-(dx * dx + dy * dy + dz * dz) * inv2s2)) / (1.0 + amp);
        if (u < f) {
            x = px;
            y = py;
            z = pz;
            ok = true;
            break;
        }
    }
    if (!ok) atomicAdd(fail, 1ull);
    double vx, vy, vz;
    maxwell3(r, vth, vx, vy, vz);
    store_loaded_mx(I, S, b, i, x, y, z, vx, vy, vz, i, I->track_ids != 0, tx_, ty_, tz_);
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <int TX, int TY, int TZ, int WO>
static hipError_t launch_tile_mx(int mode, const InstDev *insts, const WorkItem *items,
                                 int nitems, const int *parity, hipStream_t st)
{
    dim3 g(nitems), b(kBlock);
#define OHPIC_MX_CASE(M)                                                                 \
    case M:                                                                              \
        hipLaunchKernelGGL((k_particles_mixed<TX, TY, TZ, WO, M>), g, b, 0, st, insts, items, \
                           parity, nullptr, nitems);                                     \
        break;
    switch (mode) {
        OHPIC_MX_CASE(kModeFused)
        OHPIC_MX_CASE(kModeDeposit)
        OHPIC_MX_CASE(kModeGatherPush)
        OHPIC_MX_CASE(kModeBoundary)
        OHPIC_MX_CASE(kModeFusedList)
        OHPIC_MX_CASE(kModeBoundaryList)
    default:
        return hipErrorInvalidValue;
    }
#undef OHPIC_MX_CASE
    return hipGetLastError();
}

hipError_t launch_particles_mixed(int mode, int weight_order, const int T[3],
                                  const InstDev *insts, const WorkItem *items, int nitems,
                                  const int *parity, hipStream_t st)
{
#define OHPIC_TILE(A, B, C)                                                        \
    if (T[0] == A && T[1] == B && T[2] == C)                                       \
        return weight_order == 0                                                   \
                   ? launch_tile_mx<A, B, C, 0>(mode, insts, items, nitems, parity, st) \
                   : launch_tile_mx<A, B, C, 1>(mode, insts, items, nitems, parity, st);
    OHPIC_TILE(4, 4, 4)
    OHPIC_TILE(8, 8, 8)
    OHPIC_TILE(8, 8, 4)
    OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    return hipErrorInvalidValue;
}

template <int TX, int TY, int TZ, int WO>
static hipError_t launch_tile_mx_queue(const InstDev *insts, const WorkItem *items, int nitems,
                                       const int *parity, unsigned int *queue, int blocks,
                                       hipStream_t st)
{
    hipLaunchKernelGGL((k_particles_mixed<TX, TY, TZ, WO, kModeFusedQueue>), dim3(blocks),
                       dim3(kBlock), 0, st, insts, items, parity, queue, nitems);
    return hipGetLastError();
}

template <int TX, int TY, int TZ, int WO>
static int queue_blocks_per_cu_mx()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&k_particles_mixed<TX, TY, TZ, WO, kModeFusedQueue>),
            kBlock, 0) != hipSuccess) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb > 0 ? nb : 1;
}

hipError_t launch_particles_mixed_queue(int weight_order, const int T[3], const InstDev *insts,
                                        const WorkItem *items, int nitems, const int *parity,
                                        unsigned int *queue, int blocks, hipStream_t st)
{
#define OHPIC_TILE(A, B, C)                                                                   \
    if (T[0] == A && T[1] == B && T[2] == C)                                                  \
        return weight_order == 0                                                              \
                   ? launch_tile_mx_queue<A, B, C, 0>(insts, items, nitems, parity, queue, blocks, st) \
                   : launch_tile_mx_queue<A, B, C, 1>(insts, items, nitems, parity, queue, blocks, st);
    OHPIC_TILE(4, 4, 4)
    OHPIC_TILE(8, 8, 8)
    OHPIC_TILE(8, 8, 4)
    OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    return hipErrorInvalidValue;
}

int particles_mixed_queue_blocks(int weight_order, const int T[3])
{
#define OHPIC_TILE(A, B, C)                                                            \
    if (T[0] == A && T[1] == B && T[2] == C)                                           \
        return weight_order == 0 ? queue_blocks_per_cu_mx<A, B, C, 0>()                \
                                 : queue_blocks_per_cu_mx<A, B, C, 1>();
    OHPIC_TILE(4, 4, 4)
    OHPIC_TILE(8, 8, 8)
    OHPIC_TILE(8, 8, 4)
    OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    return 1;
}

hipError_t launch_bin_count_mx(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                               const int T[3], hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_count_mx, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, lbuf, n, T[0], T[1], T[2]);
    return hipGetLastError();
}

hipError_t launch_bin_scatter_mx(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                                 const int T[3], hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_scatter_mx, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, lbuf, n, T[0], T[1], T[2]);
    return hipGetLastError();
}

hipError_t launch_bins_to_list_mx(const InstDev *insts, int inst, int s, int src, int dst,
                                  const int *tiles, int t0, int ntiles, const int64_t *list_off,
                                  hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bins_to_list_mx, dim3(ntiles), dim3(256), 0, st, insts, inst, s, src,
                       dst, tiles, t0, list_off);
    return hipGetLastError();
}

hipError_t launch_cell_counts_mx(const InstDev *insts, int inst, int s, int par, int ntiles,
                                 int32_t *counts, const int T[3], hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_counts_mx, dim3(ntiles), dim3(256), 0, st, insts, inst, s, par,
                       counts, T[0], T[1], T[2]);
    return hipGetLastError();
}

hipError_t launch_load_uniform_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                                  uint64_t offset, int64_t np, double vth, const int T[3],
                                  hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load_uniform_mx, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, b, rng0, offset, np, vth, T[0], T[1], T[2]);
    return hipGetLastError();
}

hipError_t launch_load_hump_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                               uint64_t offset, uint64_t budget, int64_t np, double amp,
                               double sigma, double vth, unsigned long long *fail,
                               const int T[3], hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load_hump_mx, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, b, rng0, offset, budget, np, amp, sigma, vth, fail, T[0],
                       T[1], T[2]);
    return hipGetLastError();
}

}  // namespace ohpic
