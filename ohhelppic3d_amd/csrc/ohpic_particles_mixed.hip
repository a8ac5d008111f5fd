// ohpic_particles_mixed.hip -- mixed-precision particle kernels for gfx950.
//
// Storage per particle (28 B): an int32 cell index packing ix | iy << 10 |
// iz << 20, FP32 in-cell offsets ox, oy, oz in [0, 1) and FP32 velocities.
// Gather, Boris, drift and the deposit weights run in FP32; every deposit
// contribution is rounded into the same int64 fixed-point charge grid the
// FP64 path uses, so rho, the Poisson solve, E and the diagnostics stay FP64.
// The periodic boundary is the integer wrap of the cell index (the position
// never leaves its normalised cell + offset form).  Work decomposition,
// LDS staging, deposit blocks and re-binning are those of the FP64 kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "ohpic_internal.hpp"
#include "ohpic_kernels.hpp"
#include "ohpic_particle_common.hpp"
#include "ohpic_pcg.hpp"

namespace ohpic {

namespace {

constexpr int kCellBits = 10;
constexpr int kCellMask = (1 << kCellBits) - 1;
constexpr float kOffsetMax = 0.99999994f;  // largest float below 1

__device__ __forceinline__ int cell_pack(int x, int y, int z)
{
    return x | (y << kCellBits) | (z << (2 * kCellBits));
}

struct PMx {
    int cell;
    float ox, oy, oz, vx, vy, vz;
    int64_t id;
};

__device__ __forceinline__ PMx load_pmx(const SpecDev &S, int b, int64_t o, bool ids)
{
    PMx q;
    q.cell = ((const int32_t *)S.buf[b][0])[o];
    q.ox = ((const float *)S.buf[b][1])[o];
    q.oy = ((const float *)S.buf[b][2])[o];
    q.oz = ((const float *)S.buf[b][3])[o];
    q.vx = ((const float *)S.buf[b][4])[o];
    q.vy = ((const float *)S.buf[b][5])[o];
    q.vz = ((const float *)S.buf[b][6])[o];
    q.id = ids ? S.pid[b][o] : 0;
    return q;
}

__device__ __forceinline__ void store_pmx(const SpecDev &S, int b, int64_t o, const PMx &q,
                                          bool ids)
{
    ((int32_t *)S.buf[b][0])[o] = q.cell;
    ((float *)S.buf[b][1])[o] = q.ox;
    ((float *)S.buf[b][2])[o] = q.oy;
    ((float *)S.buf[b][3])[o] = q.oz;
    ((float *)S.buf[b][4])[o] = q.vx;
    ((float *)S.buf[b][5])[o] = q.vy;
    ((float *)S.buf[b][6])[o] = q.vz;
    if (ids) S.pid[b][o] = q.id;
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

// Move the offset back into [0, 1) and shift the cell coordinate with the
// periodic integer wrap.  Returns false when the offset is not finite.
__device__ __forceinline__ bool recell(float &o, int &c, int n)
{
    if (o >= 0.0f && o < 1.0f) return true;
    if (!(fabsf(o) < 1.0e9f)) {
        // non-finite or absurd displacement: reported by the caller as out of range
        o = 0.0f;
        return false;
    }
    const float f = floorf(o);
    int d = (int)f;
    o = o - f;
    if (o >= 1.0f) {
        o -= 1.0f;
        d += 1;
    }
    c += d;
    if (c < 0 || c >= n) c = ((c % n) + n) % n;
    return o >= 0.0f && o < 1.0f;
}

__device__ __forceinline__ void route_back_mx(const SpecDev &S, int nbuf, int t, const PMx &q,
                                              bool ids, unsigned long long *flags)
{
    const unsigned long long old = atomicAdd(&S.count[nbuf][t], 1ull << 32);
    const int f = bin_front(old), b = bin_back(old);
    const int cap = S.bin_cap[t];
    if (b + f < cap)
        store_pmx(S, nbuf, S.bin_base[t] + (cap - 1 - b), q, ids);
    else
        atomicAdd(&flags[kFlagOverflow], 1ull);
}

struct LeaveBufMx {
    float c[6][kLeaveCap];
    int32_t cell[kLeaveCap];
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
        q.cell = lb.cell[l];
        q.ox = lb.c[0][l];
        q.oy = lb.c[1][l];
        q.oz = lb.c[2][l];
        q.vx = lb.c[3][l];
        q.vy = lb.c[4][l];
        q.vz = lb.c[5][l];
        q.id = lb.id[l];
        route_back_mx(S, nbuf, lb.dest[l], q, ids, flags);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TX, int TY, int TZ, int WO, int MODE>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(OHPIC_PUSH_MIN_WAVES, 8)))
    k_particles_mixed(const InstDev *__restrict__ insts, const WorkItem *__restrict__ items,
                      const int *__restrict__ parity)
{
    constexpr bool LIST = (MODE == kModeFusedList || MODE == kModeBoundaryList);
    constexpr bool PUSH = (MODE == kModeFused || MODE == kModeGatherPush || MODE == kModeFusedList);
    constexpr bool REBIN = (MODE == kModeFused || MODE == kModeBoundary || LIST);
    constexpr bool DEP = (MODE == kModeFused || MODE == kModeDeposit || MODE == kModeFusedList);
    constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1, NE = EX * EY * EZ;
    constexpr int RX = TX + 3, RY = TY + 3, RZ = TZ + 3, NR = RX * RY * RZ;
    __shared__ float sE[PUSH ? 3 * NE : 1];
    __shared__ unsigned long long sR[DEP ? NR : 1];
    __shared__ LeaveBufMx sLeave[(REBIN && !LIST) ? kWaves : 1];
    __shared__ int sLcnt[kWaves];
    __shared__ double sRed[kWaves];
    __shared__ int sFront, sFrontLimit;

    const WorkItem it = items[blockIdx.x];
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
            const int gx = (lox + lx) % nx, gy = (loy + ly) % ny, gz = (loz + lz) % nz;
            const int64_t g = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
            sE[l] = (float)ex[g];
            sE[NE + l] = (float)ey[g];
            sE[2 * NE + l] = (float)ez[g];
        }
    }
    if (DEP)
        for (int l = tid; l < NR; l += kBlock) sR[l] = 0ull;
    if (tid < kWaves) sLcnt[tid] = 0;
    __syncthreads();

    const float dt = (float)I->dt;
    const float dt2 = 0.5f * dt;
    const float bx = (float)I->b0[0], by = (float)I->b0[1], bz = (float)I->b0[2];
    const int nbuf = par ^ 1;
    const unsigned long long lanemask_lt = (1ull << lane_id()) - 1ull;
    unsigned long long *__restrict__ racc =
        I->rho_acc[(MODE == kModeFused || MODE == kModeFusedList) ? nbuf : par];

    for (int s = 0; s < I->nspecies; ++s) {
        const SpecDev &S = I->sp[s];
        const int cap = S.bin_cap[tile];
        const unsigned long long word = S.count[par][tile];
        const int F = min(bin_front(word), cap);
        const int B = min(bin_back(word), cap - F);
        const int cnt = F + B;
        const int64_t base = S.bin_base[tile];
        int per = (cnt + it.nchunks - 1) / it.nchunks;
        per = (per + 63) & ~63;
        const int start = it.chunk * per;
        const int end = min(cnt, start + per);
        if (REBIN) {
            if (tid == 0) {
                int lim = cap;
                if (!LIST && excl && cnt > 0) {
                    const unsigned long long old =
                        atomicAdd(&S.count[nbuf][tile], (unsigned long long)cnt);
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

        PMx nextq{0, 0, 0, 0, 0, 0, 0, 0};
        if (start + tid < end)
            nextq = load_pmx(S, par, base + bin_slot(start + tid, F, B, cap), ids);
        for (int p0 = start; p0 < end; p0 += kBlock) {
            const int p = p0 + tid;
            const bool active = p < end;
            PMx q = nextq;
            if (p + kBlock < end)
                nextq = load_pmx(S, par, base + bin_slot(p + kBlock, F, B, cap), ids);
            int cx = q.cell & kCellMask, cy = (q.cell >> kCellBits) & kCellMask,
                cz = q.cell >> (2 * kCellBits);
            if (PUSH && active) {
                float wx0, wx1, wy0, wy1, wz0, wz1;
                cic_wf<WO>(q.ox, wx0, wx1);
                cic_wf<WO>(q.oy, wy0, wy1);
                cic_wf<WO>(q.oz, wz0, wz1);
                const int lx = min(max(cx - lox, 0), TX - 1);
                const int ly = min(max(cy - loy, 0), TY - 1);
                const int lz = min(max(cz - loz, 0), TZ - 1);
                const int l0 = lx + EX * (ly + EY * lz);
                float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
                const float wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const int l = l0 + i + EX * (j + EY * k);
                            const float w = wx[i] * wy[j] * wz[k];
                            e0 += sE[l] * w;
                            e1 += sE[NE + l] * w;
                            e2 += sE[2 * NE + l] * w;
                        }
                keo += (double)(q.vx * q.vx + q.vy * q.vy + q.vz * q.vz);
                const float um0 = q.vx + qm * e0 * dt2;
                const float um1 = q.vy + qm * e1 * dt2;
                const float um2 = q.vz + qm * e2 * dt2;
                const float ua0 = um0 + (um1 * t2 - um2 * t1);
                const float ua1 = um1 + (um2 * t0 - um0 * t2);
                const float ua2 = um2 + (um0 * t1 - um1 * t0);
                const float up0 = um0 + (ua1 * s2 - ua2 * s1);
                const float up1 = um1 + (ua2 * s0 - ua0 * s2);
                const float up2 = um2 + (ua0 * s1 - ua1 * s0);
                q.vx = up0 + qm * e0 * dt2;
                q.vy = up1 + qm * e1 * dt2;
                q.vz = up2 + qm * e2 * dt2;
                ken += (double)(q.vx * q.vx + q.vy * q.vy + q.vz * q.vz);
                q.ox = q.ox + q.vx * dt;
                q.oy = q.oy + q.vy * dt;
                q.oz = q.oz + q.vz * dt;
                const bool ok = recell(q.ox, cx, nx) & recell(q.oy, cy, ny) &
                                recell(q.oz, cz, nz);
                if (!ok) {
                    atomicAdd(&flags[kFlagRange], 1ull);
                    q.ox = q.oy = q.oz = 0.0f;
                }
                q.cell = cell_pack(cx, cy, cz);
            }
            if (MODE == kModeGatherPush) {
                if (active) store_pmx(S, par, base + bin_slot(p, F, B, cap), q, ids);
                continue;
            }
            const int ntile = cx / TX + ntx * (cy / TY + nty * (cz / TZ));
            if (DEP && active) {
                float wx0, wx1, wy0, wy1, wz0, wz1;
                cic_wf<WO>(q.ox, wx0, wx1);
                cic_wf<WO>(q.oy, wy0, wy1);
                cic_wf<WO>(q.oz, wz0, wz1);
                const float wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
                const int dx = wrap_index(cx - lox + 1, nx);
                const int dy = wrap_index(cy - loy + 1, ny);
                const int dz = wrap_index(cz - loz + 1, nz);
                if (dx <= TX + 1 && dy <= TY + 1 && dz <= TZ + 1) {
                    const int l0 = dx + RX * (dy + RY * dz);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int k = 0; k < 2; ++k) {
                                const long long v =
                                    fix_round((double)(wx[i] * wy[j] * wz[k]), qfix);
                                atomicAdd(&sR[l0 + i + RX * (j + RY * k)],
                                          (unsigned long long)v);
                            }
                } else {
                    atomicAdd(&flags[kFlagFast], 1ull);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int k = 0; k < 2; ++k) {
                                const int gx = (cx + i) % nx, gy = (cy + j) % ny,
                                          gz = (cz + k) % nz;
                                const long long v =
                                    fix_round((double)(wx[i] * wy[j] * wz[k]), qfix);
                                atomicAdd(&racc[gx + (int64_t)nx * (gy + (int64_t)ny * gz)],
                                          (unsigned long long)v);
                            }
                }
            }
            if (LIST) {
                const unsigned long long m = __ballot(active);
                if (m) {
                    const int leader = __ffsll((unsigned long long)m) - 1;
                    unsigned long long b = 0;
                    if (lane_id() == leader)
                        b = atomicAdd(&S.count[nbuf][0], (unsigned long long)__popcll(m));
                    b = __shfl(b, leader, 64);
                    if (active)
                        store_pmx(S, nbuf, (int64_t)b + __popcll(m & lanemask_lt), q, ids);
                }
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
                            atomicAdd(&S.count[nbuf][tile], (unsigned long long)nst);
                        b = bin_front(old);
                        lim = cap - bin_back(old);
                    }
                    b = __shfl(b, leader, 64);
                    lim = __shfl(lim, leader, 64);
                    if (stay) {
                        const int slot = b + (int)__popcll(m & lanemask_lt);
                        if (slot < lim)
                            store_pmx(S, nbuf, base + slot, q, ids);
                        else
                            atomicAdd(&flags[kFlagOverflow], 1ull);
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
                            lb.cell[slot] = q.cell;
                            lb.c[0][slot] = q.ox;
                            lb.c[1][slot] = q.oy;
                            lb.c[2][slot] = q.oz;
                            lb.c[3][slot] = q.vx;
                            lb.c[4][slot] = q.vy;
                            lb.c[5][slot] = q.vz;
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
                atomicAdd(&S.count[nbuf][tile],
                          (unsigned long long)(-(long long)(cnt - sFront)));
            __syncthreads();
        }
        if (PUSH) {
            const double so = block_sum(keo, sRed);
            const double sn = block_sum(ken, sRed);
            if (tid == 0) {
                I->ke_partial[(int64_t)it.slot * kKeSlots + s] = 0.5 * S.mass * so;
                I->ke_partial[(int64_t)it.slot * kKeSlots + kMaxSpecies + s] = 0.5 * S.mass * sn;
            }
        }
    }

    if (DEP) {
        __syncthreads();
        for (int l = tid; l < NR; l += kBlock) {
            const unsigned long long v = sR[l];
            if (v == 0ull) continue;
            const int lx = l % RX, ly = (l / RX) % RY, lz = l / (RX * RY);
            const int gx = ((lox - 1 + lx) % nx + nx) % nx;
            const int gy = ((loy - 1 + ly) % ny + ny) % ny;
            const int gz = ((loz - 1 + lz) % nz + nz) % nz;
            atomicAdd(&racc[gx + (int64_t)nx * (gy + (int64_t)ny * gz)], v);
        }
    }
}

// ---------------------------------------------------------------- binning
__device__ __forceinline__ bool valid_mx(const InstDev *I, const PMx &q)
{
    const int cx = q.cell & kCellMask, cy = (q.cell >> kCellBits) & kCellMask,
              cz = q.cell >> (2 * kCellBits);
    return q.cell >= 0 && cx < I->n[0] && cy < I->n[1] && cz < I->n[2] && q.ox >= 0.0f &&
           q.ox < 1.0f && q.oy >= 0.0f && q.oy < 1.0f && q.oz >= 0.0f && q.oz < 1.0f;
}

__device__ __forceinline__ int tile_of_mx(const InstDev *I, int cell, int tx_, int ty_, int tz_)
{
    const int cx = cell & kCellMask, cy = (cell >> kCellBits) & kCellMask,
              cz = cell >> (2 * kCellBits);
    return cx / tx_ + I->nt[0] * (cy / ty_ + I->nt[1] * (cz / tz_));
}

__global__ void k_bin_count_mx(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                               int64_t n, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PMx q = load_pmx(S, lbuf, i, false);
    if (!valid_mx(I, q)) {
        atomicAdd(&I->flags[kFlagRange], 1ull);
        return;
    }
    atomicAdd(&S.count[lbuf ^ 1][tile_of_mx(I, q.cell, tx_, ty_, tz_)], 1ull);
}

__global__ void k_bin_scatter_mx(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                                 int64_t n, int tx_, int ty_, int tz_)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ids = I->track_ids != 0;
    const PMx q = load_pmx(S, lbuf, i, ids);
    if (!valid_mx(I, q)) return;
    const int t = tile_of_mx(I, q.cell, tx_, ty_, tz_);
    const unsigned long long old = atomicAdd(&S.count[lbuf ^ 1][t], 1ull);
    const int slot = bin_front(old);
    if (slot < S.bin_cap[t])
        store_pmx(S, lbuf ^ 1, S.bin_base[t] + slot, q, ids);
    else
        atomicAdd(&I->flags[kFlagOverflow], 1ull);
}

__global__ void k_bins_to_list_mx(const InstDev *__restrict__ insts, int inst, int s, int src,
                                  int dst, const int64_t *__restrict__ list_off)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int t = blockIdx.x;
    const unsigned long long w = S.count[src][t];
    const int cap = S.bin_cap[t];
    const int F = min(bin_front(w), cap);
    const int B = min(bin_back(w), cap - F);
    const int64_t ib = S.bin_base[t];
    const int64_t ob = list_off[t];
    const bool ids = I->track_ids != 0;
    for (int v = threadIdx.x; v < F + B; v += blockDim.x) {
        const int64_t o = ib + bin_slot(v, F, B, cap);
#pragma unroll
        for (int c = 0; c < 7; ++c)
            ((uint32_t *)S.buf[dst][c])[ob + v] = ((const uint32_t *)S.buf[src][c])[o];
        if (ids) S.pid[dst][ob + v] = S.pid[src][o];
    }
}

__global__ void k_cell_counts_mx(const InstDev *__restrict__ insts, int inst, int s, int par,
                                 int32_t *__restrict__ counts)
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
    for (int v = threadIdx.x; v < F + B; v += blockDim.x) {
        const int cell = ((const int32_t *)S.buf[par][0])[b + bin_slot(v, F, B, cap)];
        const int cx = cell & kCellMask, cy = (cell >> kCellBits) & kCellMask,
                  cz = cell >> (2 * kCellBits);
        atomicAdd(&counts[cx + (int64_t)nx * (cy + (int64_t)ny * cz)], 1);
    }
}

// ---------------------------------------------------------------- loaders
__device__ __forceinline__ void store_loaded_mx(const SpecDev &S, int b, int64_t i, double x,
                                                double y, double z, double vx, double vy,
                                                double vz, const int n[3], int64_t id, bool ids)
{
    const double p[3] = {x, y, z};
    int c[3];
    float o[3];
    for (int a = 0; a < 3; ++a) {
        // keep the FP64 cell (int() truncation): an offset that rounds up to
        // 1 is clamped below it, so the particle stays on the same side of
        // the cell face as in FP64 (the mirrored CIC weights jump there)
        c[a] = (int)p[a];
        o[a] = fminf((float)(p[a] - (double)c[a]), kOffsetMax);
    }
    PMx q{cell_pack(c[0], c[1], c[2]), o[0], o[1], o[2], (float)vx, (float)vy, (float)vz, id};
    store_pmx(S, b, i, q, ids);
}

__global__ void k_load_uniform_mx(const InstDev *__restrict__ insts, int inst, int s, int b,
                                  Pcg32Dev rng0, uint64_t offset, int64_t np, double vth)
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
    store_loaded_mx(S, b, i, x, y, z, vx, vy, vz, I->n, i, I->track_ids != 0);
}

__global__ void k_load_hump_mx(const InstDev *__restrict__ insts, int inst, int s, int b,
                               Pcg32Dev rng0, uint64_t offset, uint64_t budget, int64_t np,
                               double amp, double sigma, double vth,
                               unsigned long long *__restrict__ fail)
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
    store_loaded_mx(S, b, i, x, y, z, vx, vy, vz, I->n, i, I->track_ids != 0);
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
                           parity);                                                      \
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
                                  int ntiles, const int64_t *list_off, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bins_to_list_mx, dim3(ntiles), dim3(256), 0, st, insts, inst, s, src,
                       dst, list_off);
    return hipGetLastError();
}

hipError_t launch_cell_counts_mx(const InstDev *insts, int inst, int s, int par, int ntiles,
                                 int32_t *counts, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_counts_mx, dim3(ntiles), dim3(256), 0, st, insts, inst, s, par,
                       counts);
    return hipGetLastError();
}

hipError_t launch_load_uniform_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                                  uint64_t offset, int64_t np, double vth, hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load_uniform_mx, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, b, rng0, offset, np, vth);
    return hipGetLastError();
}

hipError_t launch_load_hump_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                               uint64_t offset, uint64_t budget, int64_t np, double amp,
                               double sigma, double vth, unsigned long long *fail,
                               hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load_hump_mx, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                       insts, inst, s, b, rng0, offset, budget, np, amp, sigma, vth, fail);
    return hipGetLastError();
}

}  // namespace ohpic
