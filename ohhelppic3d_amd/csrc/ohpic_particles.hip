// ohpic_particles.hip -- particle kernels for gfx950 (FP64 storage).
//
// One workgroup processes one work item = (instance, tile, chunk): it stages
// the node-centred E of the tile plus a one-node halo in LDS, walks its slice
// of every species' bin, and for each particle
//   gathers E with CIC weights   (linear_interpolator.f90:22-69),
//   runs the Boris rotation + drift (particle_mover_boris.f90:21-64),
//   applies the periodic wrap twice (periodic_particle_boundary.f90:42-53,
//                                    ohhelppic3d.f90:353-355),
//   deposits q*W(x_new) into an int64 fixed-point LDS copy of the tile's
//   nodes plus halo (linear_scatter.f90:22-55), and
//   re-bins it (the replacement of OhHelp's correct_particle/transbound):
//   a particle that stays in its tile goes to the front run of the tile's
//   bin in the other buffer (slot from an LDS counter), a particle that left
//   goes through a per-wave LDS buffer to the back run of its new tile's bin
//   (batched 64-bit atomics on the bin's front|back word).
// The LDS charge block is flushed with one global 64-bit atomic per node.
// Integer accumulation makes rho independent of particle order and timing.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ohpic_internal.hpp"
#include "ohpic_kernels.hpp"
#include "ohpic_particle_common.hpp"

namespace ohpic {

namespace {

struct P64 {
    double x, y, z, vx, vy, vz;
    int64_t id;
};

__device__ __forceinline__ void store_p64(const SpecDev &S, int nbuf, int64_t o, const P64 &p,
                                          bool ids)
{
    sstore((double *)S.buf[nbuf][0], o, p.x);
    sstore((double *)S.buf[nbuf][1], o, p.y);
    sstore((double *)S.buf[nbuf][2], o, p.z);
    sstore((double *)S.buf[nbuf][3], o, p.vx);
    sstore((double *)S.buf[nbuf][4], o, p.vy);
    sstore((double *)S.buf[nbuf][5], o, p.vz);
    if (ids) gstore(S.pid[nbuf], o, p.id);
}

__device__ __forceinline__ P64 load_p64(const SpecDev &S, int b, int64_t o, bool ids)
{
    P64 q;
    q.x = sload((const double *)S.buf[b][0], o);
    q.y = sload((const double *)S.buf[b][1], o);
    q.z = sload((const double *)S.buf[b][2], o);
    q.vx = sload((const double *)S.buf[b][3], o);
    q.vy = sload((const double *)S.buf[b][4], o);
    q.vz = sload((const double *)S.buf[b][5], o);
    q.id = ids ? gload(S.pid[b], o) : 0;
    return q;
}

// The bin of one tile in one buffer: array bases offset to the bin's first
// element (uniform per workgroup: scalar registers) and particles addressed
// by a 32-bit byte offset (bins hold < 2^29 particles), so per-particle loads
// and stores need no 64-bit address arithmetic.
struct BinView64 {
    const char *c[6];
    const int64_t *id;
};

__device__ __forceinline__ BinView64 bin_view64(const SpecDev &S, int b, int64_t base, bool ids)
{
    BinView64 r;
#pragma unroll
    for (int k = 0; k < 6; ++k) r.c[k] = (const char *)((const double *)S.buf[b][k] + base);
    r.id = ids ? S.pid[b] + base : nullptr;
    return r;
}

__device__ __forceinline__ P64 load_bin64(const BinView64 &bv, uint32_t slot, bool ids)
{
    const uint32_t off = slot * 8u;
    P64 q;
    q.x = sload((const double *)(bv.c[0] + off), 0);
    q.y = sload((const double *)(bv.c[1] + off), 0);
    q.z = sload((const double *)(bv.c[2] + off), 0);
    q.vx = sload((const double *)(bv.c[3] + off), 0);
    q.vy = sload((const double *)(bv.c[4] + off), 0);
    q.vz = sload((const double *)(bv.c[5] + off), 0);
    q.id = ids ? gload(bv.id, (int64_t)slot) : 0;
    return q;
}

__device__ __forceinline__ void store_bin64(const BinView64 &bv, uint32_t slot, const P64 &q,
                                            bool ids)
{
    const uint32_t off = slot * 8u;
    sstore((double *)(bv.c[0] + off), 0, q.x);
    sstore((double *)(bv.c[1] + off), 0, q.y);
    sstore((double *)(bv.c[2] + off), 0, q.z);
    sstore((double *)(bv.c[3] + off), 0, q.vx);
    sstore((double *)(bv.c[4] + off), 0, q.vy);
    sstore((double *)(bv.c[5] + off), 0, q.vz);
    if (ids) gstore(const_cast<int64_t *>(bv.id), (int64_t)slot, q.id);
}

// Append one particle to the back run of bin t (slot from the packed word).
__device__ __forceinline__ void route_back(const SpecDev &S, int nbuf, int t, const P64 &q,
                                           bool ids, unsigned long long *flags)
{
    const unsigned long long old = gatomic_add(&S.count[nbuf][t], 1ull << 32);
    const int f = bin_front(old), b = bin_back(old);
    const int cap = gload(S.bin_cap, t);
    if (b + f < cap)
        store_p64(S, nbuf, gload(S.bin_base, t) + (cap - 1 - b), q, ids);
    else
        gatomic_add(&flags[kFlagOverflow], 1ull);
}

struct LeaveBuf {
    double c[6][kLeaveCap];
    int64_t id[kLeaveCap];
    int32_t dest[kLeaveCap];
};

// Route the first `cnt` buffered leavers of this wave (lane l takes entry l).
__device__ __forceinline__ void flush_leavers(LeaveBuf &lb, int cnt, const SpecDev &S, int nbuf,
                                              bool ids, unsigned long long *flags)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int l = lane_id();
    if (l < cnt) {
        P64 q;
        q.x = lb.c[0][l];
        q.y = lb.c[1][l];
        q.z = lb.c[2][l];
        q.vx = lb.c[3][l];
        q.vy = lb.c[4][l];
        q.vz = lb.c[5][l];
        q.id = lb.id[l];
        route_back(S, nbuf, lb.dest[l], q, ids, flags);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TX, int TY, int TZ, int WO, int MODE_>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(OHPIC_PUSH_MIN_WAVES, 8)))
    k_particles_fp64(const InstDev *__restrict__ insts, const WorkItem *__restrict__ items,
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
    __shared__ double sE[PUSH ? 3 * NE : 1];
    __shared__ unsigned long long sR[DEP ? NR : 1];
    __shared__ LeaveBuf sLeave[(REBIN && !LIST) ? kWaves : 1];
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
        const bool excl = it.nchunks == 1;  // this workgroup alone fills the tile's front runs
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
                sE[l] = gload(ex, g);
                sE[NE + l] = gload(ey, g);
                sE[2 * NE + l] = gload(ez, g);
            }
        }
        if (DEP)
            for (int l = tid; l < NR; l += kBlock) sR[l] = 0ull;
        if (tid < kWaves) sLcnt[tid] = 0;
        __syncthreads();

        const double dt = I->dt;
        const double dt2 = dt * 0.5;
        const double bx = I->b0[0], by = I->b0[1], bz = I->b0[2];
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
                        // reserve the whole front run: arrivals from neighbours stay behind it
                        const unsigned long long old =
                            gatomic_add(&S.count[nbuf][tile], (unsigned long long)cnt);
                        lim = cap - bin_back(old);
                    }
                    sFront = 0;
                    sFrontLimit = lim;
                }
                __syncthreads();
            }
            const double qm = S.qm;
            const double qfix = S.qfix;
            // Boris constants for the uniform B (t = qm B dt/2, s = 2t/(1+|t|^2))
            const double t0 = qm * bx * dt2, t1 = qm * by * dt2, t2 = qm * bz * dt2;
            const double tt = t0 * t0 + t1 * t1 + t2 * t2;
            const double s0 = 2 * t0 / (1 + tt), s1 = 2 * t1 / (1 + tt), s2 = 2 * t2 / (1 + tt);
            double keo = 0.0, ken = 0.0;

            const BinView64 cur = bin_view64(S, par, base, ids);
            const BinView64 nxt = bin_view64(S, nbuf, base, ids);
            P64 nextq{0, 0, 0, 0, 0, 0, 0};
            if (start + tid < end)
                nextq = load_bin64(cur, (uint32_t)bin_slot(start + tid, F, B, cap), ids);
            for (int p0 = start; p0 < end; p0 += kBlock) {
                const int p = p0 + tid;
                const bool active = p < end;
                P64 q = nextq;
                // prefetch the next iteration's particle while this one is processed
                if (p + kBlock < end)
                    nextq = load_bin64(cur, (uint32_t)bin_slot(p + kBlock, F, B, cap), ids);
                if (PUSH && active) {
                    const int ix = (int)q.x, iy = (int)q.y, iz = (int)q.z;
                    double wx0, wx1, wy0, wy1, wz0, wz1;
                    cic_w<WO>(q.x - (double)ix, wx0, wx1);
                    cic_w<WO>(q.y - (double)iy, wy0, wy1);
                    cic_w<WO>(q.z - (double)iz, wz0, wz1);
                    // binned particles lie in the tile; the clamp only guards LDS
                    const int lx = min(max(ix - lox, 0), TX - 1);
                    const int ly = min(max(iy - loy, 0), TY - 1);
                    const int lz = min(max(iz - loz, 0), TZ - 1);
                    const int l0 = lx + EX * (ly + EY * lz);
                    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
                    const double wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int k = 0; k < 2; ++k) {
                                const int l = l0 + i + EX * (j + EY * k);
                                const double w = wx[i] * wy[j] * wz[k];
                                e0 += sE[l] * w;
                                e1 += sE[NE + l] * w;
                                e2 += sE[2 * NE + l] * w;
                            }
                    keo += q.vx * q.vx + q.vy * q.vy + q.vz * q.vz;
                    // Boris (particle_mover_boris.f90:42-62)
                    const double um0 = q.vx + qm * e0 * dt2;
                    const double um1 = q.vy + qm * e1 * dt2;
                    const double um2 = q.vz + qm * e2 * dt2;
                    const double ua0 = um0 + (um1 * t2 - um2 * t1);
                    const double ua1 = um1 + (um2 * t0 - um0 * t2);
                    const double ua2 = um2 + (um0 * t1 - um1 * t0);
                    const double up0 = um0 + (ua1 * s2 - ua2 * s1);
                    const double up1 = um1 + (ua2 * s0 - ua0 * s2);
                    const double up2 = um2 + (ua0 * s1 - ua1 * s0);
                    q.vx = up0 + qm * e0 * dt2;
                    q.vy = up1 + qm * e1 * dt2;
                    q.vz = up2 + qm * e2 * dt2;
                    ken += q.vx * q.vx + q.vy * q.vy + q.vz * q.vz;
                    q.x = q.x + q.vx * dt;
                    q.y = q.y + q.vy * dt;
                    q.z = q.z + q.vz * dt;
                }
                if (MODE == kModeGatherPush) {
                    if (active) store_bin64(cur, (uint32_t)bin_slot(p, F, B, cap), q, ids);
                    continue;
                }
                int ntile = tile;
                if (REBIN && active) {
                    q.x = apply_periodic(apply_periodic(q.x, nx), nx);
                    q.y = apply_periodic(apply_periodic(q.y, ny), ny);
                    q.z = apply_periodic(apply_periodic(q.z, nz), nz);
                    if (!(q.x >= 0.0 && q.x < nx && q.y >= 0.0 && q.y < ny && q.z >= 0.0 &&
                          q.z < nz)) {
                        gatomic_add(&flags[kFlagRange], 1ull);
                        q.x = fmin(fmax(q.x - floor(q.x / nx) * nx, 0.0), nx - 1e-9);
                        q.y = fmin(fmax(q.y - floor(q.y / ny) * ny, 0.0), ny - 1e-9);
                        q.z = fmin(fmax(q.z - floor(q.z / nz) * nz, 0.0), nz - 1e-9);
                    }
                    const int cxi = (int)q.x, cyi = (int)q.y, czi = (int)q.z;
                    ntile = cxi / TX + ntx * (cyi / TY + nty * (czi / TZ));
                }
                if (DEP && active) {
                    const int ix = (int)q.x, iy = (int)q.y, iz = (int)q.z;
                    double wx0, wx1, wy0, wy1, wz0, wz1;
                    cic_w<WO>(q.x - (double)ix, wx0, wx1);
                    cic_w<WO>(q.y - (double)iy, wy0, wy1);
                    cic_w<WO>(q.z - (double)iz, wz0, wz1);
                    const double wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
                    const int dx = wrap_index(ix - lox + 1, nx);
                    const int dy = wrap_index(iy - loy + 1, ny);
                    const int dz = wrap_index(iz - loz + 1, nz);
                    if (dx <= TX + 1 && dy <= TY + 1 && dz <= TZ + 1) {
                        const int l0 = dx + RX * (dy + RY * dz);
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 2; ++j)
#pragma unroll
                                for (int k = 0; k < 2; ++k) {
                                    const long long v = fix_round(wx[i] * wy[j] * wz[k], qfix);
                                    atomicAdd(&sR[l0 + i + RX * (j + RY * k)],
                                              (unsigned long long)v);
                                }
                    } else {
                        gatomic_add(&flags[kFlagFast], 1ull);
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 2; ++j)
#pragma unroll
                                for (int k = 0; k < 2; ++k) {
                                    const int gx = (ix + i) % nx, gy = (iy + j) % ny,
                                              gz = (iz + k) % nz;
                                    const long long v = fix_round(wx[i] * wy[j] * wz[k], qfix);
                                    gatomic_add(&racc[gx + (int64_t)nx * (gy + (int64_t)ny * gz)],
                                              (unsigned long long)v);
                                }
                    }
                }
                if (LIST) {
                    // retry path: every particle to the species' contiguous
                    // list, in bin order inside this item's reserved range
                    if (active) store_p64(S, nbuf, sListBase + (p - start), q, ids);
                } else if (REBIN) {
                    // particles that stayed: front run of this tile's bin
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
                                store_bin64(nxt, (uint32_t)slot, q, ids);
                            else
                                gatomic_add(&flags[kFlagOverflow], 1ull);
                        }
                    }
                    // particles that left: per-wave LDS buffer, routed in batches
                    const bool leave = active && ntile != tile;
                    const unsigned long long lm = __ballot(leave);
                    if (lm) {
                        const int nl = (int)__popcll(lm);
                        int c = sLcnt[wv];
                        if (c + nl > kLeaveCap) {
                            flush_leavers(sLeave[wv], c, S, nbuf, ids, flags);
                            c = 0;
                        }
                        if (nl <= kLeaveCap) {
                            if (leave) {
                                const int slot = c + (int)__popcll(lm & lanemask_lt);
                                LeaveBuf &lb = sLeave[wv];
                                lb.c[0][slot] = q.x;
                                lb.c[1][slot] = q.y;
                                lb.c[2][slot] = q.z;
                                lb.c[3][slot] = q.vx;
                                lb.c[4][slot] = q.vy;
                                lb.c[5][slot] = q.vz;
                                lb.id[slot] = q.id;
                                lb.dest[slot] = ntile;
                            }
                            c += nl;
                        } else if (leave) {
                            route_back(S, nbuf, ntile, q, ids, flags);
                        }
                        if (lane_id() == 0) sLcnt[wv] = c;
                    }
                }
            }
            if (REBIN && !LIST) {
                flush_leavers(sLeave[wv], sLcnt[wv], S, nbuf, ids, flags);
                __syncthreads();
                if (tid < kWaves) sLcnt[tid] = 0;
                if (tid == 0 && excl && cnt > sFront) {
                    // release the reserved but unused part of the front run
                    gatomic_add(&S.count[nbuf][tile],
                              (unsigned long long)(-(long long)(cnt - sFront)));
                }
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
__device__ __forceinline__ bool in_box(const InstDev *I, double x, double y, double z)
{
    return x >= 0.0 && x < I->n[0] && y >= 0.0 && y < I->n[1] && z >= 0.0 && z < I->n[2];
}

// Count a contiguous list [0, n) of buffer lbuf into the bin words of buffer
// lbuf^1 (front runs).  wrap != 0 applies the periodic wrap twice first.
__global__ void k_bin_count(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                            int64_t n, int tx_, int ty_, int tz_, int wrap)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double x = 0.0, y = 0.0, z = 0.0;
    if (live) {
        x = gload((const double *)S.buf[lbuf][0], i);
        y = gload((const double *)S.buf[lbuf][1], i);
        z = gload((const double *)S.buf[lbuf][2], i);
        if (wrap) {
            x = apply_periodic(apply_periodic(x, I->n[0]), I->n[0]);
            y = apply_periodic(apply_periodic(y, I->n[1]), I->n[1]);
            z = apply_periodic(apply_periodic(z, I->n[2]), I->n[2]);
        }
    }
    const bool inb = live && in_box(I, x, y, z);
    if (live && !inb) gatomic_add(&I->flags[kFlagRange], 1ull);
    const int t =
        inb ? ((int)x) / tx_ + I->nt[0] * (((int)y) / ty_ + I->nt[1] * (((int)z) / tz_)) : 0;
    (void)wave_counter_add(S.count[lbuf ^ 1], t, inb);
}

// Scatter the list into the front runs of buffer lbuf^1 (words zeroed on
// entry; caps already sized from k_bin_count).
__global__ void k_bin_scatter(const InstDev *__restrict__ insts, int inst, int s, int lbuf,
                              int64_t n, int tx_, int ty_, int tz_, int wrap)
{
    const InstDev *I = insts + inst;
    const SpecDev &S = I->sp[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const bool ids = I->track_ids != 0;
    P64 q{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    if (live) {
        q = load_p64(S, lbuf, i, ids);
        if (wrap) {
            q.x = apply_periodic(apply_periodic(q.x, I->n[0]), I->n[0]);
            q.y = apply_periodic(apply_periodic(q.y, I->n[1]), I->n[1]);
            q.z = apply_periodic(apply_periodic(q.z, I->n[2]), I->n[2]);
        }
    }
    const bool inb = live && in_box(I, q.x, q.y, q.z);
    const int t = inb ? ((int)q.x) / tx_ + I->nt[0] * (((int)q.y) / ty_ + I->nt[1] * (((int)q.z) / tz_))
                      : 0;
    const unsigned long long old = wave_counter_add(S.count[lbuf ^ 1], t, inb);
    if (!inb) return;
    const int slot = bin_front(old);
    if (slot < gload(S.bin_cap, t))
        store_p64(S, lbuf ^ 1, gload(S.bin_base, t) + slot, q, ids);
    else
        gatomic_add(&I->flags[kFlagOverflow], 1ull);
}

// Compact both runs of bins tiles[i] (tiles null: t0 + i) of buffer src into
// contiguous lists of buffer dst at list_off[i].  One workgroup per bin.  With
// src == dst each list must lie outside every bin's current slots.
__global__ void k_bins_to_list(const InstDev *__restrict__ insts, int inst, int s, int src,
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
        const int64_t o = ib + bin_slot(v, F, B, cap);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            ((double *)S.buf[dst][c])[ob + v] = ((const double *)S.buf[src][c])[o];
        if (ids) S.pid[dst][ob + v] = S.pid[src][o];
    }
}

__global__ void k_cell_counts(const InstDev *__restrict__ insts, int inst, int s, int par,
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
        const int64_t o = b + bin_slot(v, F, B, cap);
        const double x = ((const double *)S.buf[par][0])[o];
        const double y = ((const double *)S.buf[par][1])[o];
        const double z = ((const double *)S.buf[par][2])[o];
        const int64_t c = (int)x + (int64_t)nx * ((int)y + (int64_t)ny * (int)z);
        atomicAdd(&counts[c], 1);
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <int TX, int TY, int TZ, int WO>
static hipError_t launch_tile(int mode, const InstDev *insts, const WorkItem *items,
                              int nitems, const int *parity, hipStream_t st)
{
    dim3 g(nitems), b(kBlock);
#define OHPIC_FP_CASE(M)                                                                     \
    case M:                                                                                  \
        hipLaunchKernelGGL((k_particles_fp64<TX, TY, TZ, WO, M>), g, b, 0, st, insts, items, \
                           parity, nullptr, nitems);                                         \
        break;
    switch (mode) {
        OHPIC_FP_CASE(kModeFused)
        OHPIC_FP_CASE(kModeDeposit)
        OHPIC_FP_CASE(kModeGatherPush)
        OHPIC_FP_CASE(kModeBoundary)
        OHPIC_FP_CASE(kModeFusedList)
        OHPIC_FP_CASE(kModeBoundaryList)
    default:
        return hipErrorInvalidValue;
    }
#undef OHPIC_FP_CASE
    return hipGetLastError();
}

template <int TX, int TY, int TZ, int WO>
static hipError_t launch_tile_queue(const InstDev *insts, const WorkItem *items, int nitems,
                                    const int *parity, unsigned int *queue, int blocks,
                                    hipStream_t st)
{
    hipLaunchKernelGGL((k_particles_fp64<TX, TY, TZ, WO, kModeFusedQueue>), dim3(blocks),
                       dim3(kBlock), 0, st, insts, items, parity, queue, nitems);
    return hipGetLastError();
}

template <int TX, int TY, int TZ, int WO>
static int queue_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&k_particles_fp64<TX, TY, TZ, WO, kModeFusedQueue>),
            kBlock, 0) != hipSuccess) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb > 0 ? nb : 1;
}

bool tile_supported(int tx, int ty, int tz)
{
    return (tx == ty && ty == tz && (tx == 4 || tx == 8)) || (tx == 8 && ty == 8 && tz == 4) ||
           (tx == 16 && ty == 8 && tz == 4);
}

hipError_t launch_particles(int mode, int precision, int weight_order, const int T[3],
                            const InstDev *insts, const WorkItem *items, int nitems, const int *parity,
                            hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (precision == 1)
        return launch_particles_mixed(mode, weight_order, T, insts, items, nitems, parity, st);
    if (precision != 0) return hipErrorInvalidValue;
#define OHPIC_TILE(A, B, C)                                                     \
    if (T[0] == A && T[1] == B && T[2] == C)                                    \
        return weight_order == 0                                                \
                   ? launch_tile<A, B, C, 0>(mode, insts, items, nitems, parity, st) \
                   : launch_tile<A, B, C, 1>(mode, insts, items, nitems, parity, st);
    OHPIC_TILE(4, 4, 4)
    OHPIC_TILE(8, 8, 8)
    OHPIC_TILE(8, 8, 4)
    OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    return hipErrorInvalidValue;
}

hipError_t launch_particles_queue(int precision, int weight_order, const int T[3],
                                  const InstDev *insts, const WorkItem *items, int nitems,
                                  const int *parity, unsigned int *queue, int blocks,
                                  hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (blocks > nitems) blocks = nitems;
    if (blocks < 1) blocks = 1;
    if (precision == 1)
        return launch_particles_mixed_queue(weight_order, T, insts, items, nitems, parity, queue,
                                            blocks, st);
    if (precision != 0) return hipErrorInvalidValue;
#define OHPIC_TILE(A, B, C)                                                                    \
    if (T[0] == A && T[1] == B && T[2] == C)                                                   \
        return weight_order == 0                                                               \
                   ? launch_tile_queue<A, B, C, 0>(insts, items, nitems, parity, queue, blocks, st) \
                   : launch_tile_queue<A, B, C, 1>(insts, items, nitems, parity, queue, blocks, st);
    OHPIC_TILE(4, 4, 4)
    OHPIC_TILE(8, 8, 8)
    OHPIC_TILE(8, 8, 4)
    OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    return hipErrorInvalidValue;
}

int particles_queue_blocks(int precision, int weight_order, const int T[3], int device)
{
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        cus = prop.multiProcessorCount;
    else
        (void)hipGetLastError();
    int per = 1;
    if (precision == 1) {
        per = particles_mixed_queue_blocks(weight_order, T);
    } else {
#define OHPIC_TILE(A, B, C)                                                              \
    if (T[0] == A && T[1] == B && T[2] == C)                                             \
        per = weight_order == 0 ? queue_blocks_per_cu<A, B, C, 0>() : queue_blocks_per_cu<A, B, C, 1>();
        OHPIC_TILE(4, 4, 4)
        OHPIC_TILE(8, 8, 8)
        OHPIC_TILE(8, 8, 4)
        OHPIC_TILE(16, 8, 4)
#undef OHPIC_TILE
    }
    return per * cus;
}

hipError_t launch_bin_count(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                            const int T[3], int wrap, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_bin_count, dim3((unsigned)blocks), dim3(256), 0, st, insts, inst, s,
                       lbuf, n, T[0], T[1], T[2], wrap);
    return hipGetLastError();
}

hipError_t launch_bin_scatter(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                              const int T[3], int wrap, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)blocks), dim3(256), 0, st, insts, inst, s,
                       lbuf, n, T[0], T[1], T[2], wrap);
    return hipGetLastError();
}

// Bins tiles[i] of species s take base[i] and cap[i], and used[i] particles
// as one front run of buffer par (after their particles moved there).
__global__ void k_set_bins(const InstDev *__restrict__ insts, int inst, int s, int par,
                           const int *__restrict__ tiles, const int64_t *__restrict__ base,
                           const int32_t *__restrict__ cap, const int64_t *__restrict__ used,
                           int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SpecDev &S = insts[inst].sp[s];
    const int t = tiles[i];
    S.bin_base[t] = base[i];
    S.bin_cap[t] = cap[i];
    S.count[par][t] = (unsigned long long)used[i];
}

hipError_t launch_set_bins(const InstDev *insts, int inst, int s, int par, const int *tiles,
                           const int64_t *base, const int32_t *cap, const int64_t *used, int n,
                           hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_bins, dim3((n + 255) / 256), dim3(256), 0, st, insts, inst, s, par,
                       tiles, base, cap, used, n);
    return hipGetLastError();
}

hipError_t launch_bins_to_list(const InstDev *insts, int inst, int s, int src, int dst,
                               const int *tiles, int t0, int ntiles, const int64_t *list_off,
                               hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bins_to_list, dim3(ntiles), dim3(256), 0, st, insts, inst, s, src, dst,
                       tiles, t0, list_off);
    return hipGetLastError();
}

hipError_t launch_cell_counts(const InstDev *insts, int inst, int s, int par, int ntiles,
                              int32_t *counts, hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_counts, dim3(ntiles), dim3(256), 0, st, insts, inst, s, par,
                       counts);
    return hipGetLastError();
}

}  // namespace ohpic
