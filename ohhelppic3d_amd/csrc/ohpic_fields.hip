// ohpic_fields.hip -- grid kernels (fills, charge conversion, node-centred E
// with energy diagnostics) and the seeded particle loaders for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "ohpic_internal.hpp"
#include "ohpic_kernels.hpp"
#include "ohpic_pcg.hpp"

namespace ohpic {

namespace {

template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t n, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// E_a(i) = 0.5*((phi(i-e_a) - phi(i)) + (phi(i) - phi(i+e_a))): the edge
// difference of poisson_field_solver.f90:139-144 averaged onto nodes as at
// :165-170, periodic.  blockIdx.y = instance; each block walks whole x rows
// (32-bit row indexing) and writes its partial sums of rho*phi, |E|^2 and
// rho to e_partial[3*block]; k_reduce_e_b adds them in block order.  With an
// even nx every thread takes V = 2 neighbouring x points: 16-byte loads of
// the centre, y and z neighbour pairs and of rho, 8-byte loads of the two x
// neighbours outside the pair, 16-byte stores.
typedef double d2v __attribute__((ext_vector_type(2)));

// E is written with non-temporal stores: 134 us against 157 us for plain
// stores at 256^3 (the next reader, the particle kernel, finds the grid in
// HBM either way at that size; 5.3 us for both at 64^3).
template <typename T>
__device__ __forceinline__ void efield_store(gmem_t<T> *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

template <int V>
__device__ __forceinline__ void efield_rows(const InstDev *__restrict__ I, double &we, double &we2,
                                            double &qn)
{
    const gmem_t<double> *__restrict__ phi = as_global(I->phi);
    const gmem_t<double> *__restrict__ rho = as_global(I->rho);
    gmem_t<double> *__restrict__ ex = as_global(I->e[0]);
    gmem_t<double> *__restrict__ ey = as_global(I->e[1]);
    gmem_t<double> *__restrict__ ez = as_global(I->e[2]);
    const int nx = I->n[0], ny = I->n[1], nz = I->n[2];
    const int rows = ny * nz;
    const int tpr = nx / V;
    const int rpb = efield_rows_per_block(nx);
    const int nblk = efield_blocks_of(nx, ny, nz);
    // thread -> (row within the block's rows, first x); x steps by nx when
    // rows share the block, by V * the block size when one row spans it
    const int tsub = rpb > 1 ? (int)threadIdx.x / tpr : 0;
    const int tx0 = (rpb > 1 ? (int)threadIdx.x - tsub * tpr : (int)threadIdx.x) * V;
    const int xstep = rpb > 1 ? nx : (int)blockDim.x * V;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    // XCD-aware rows: with a block count divisible by 8, XCD k (blocks
    // k, k+8, ...) walks the contiguous rows [k, k+1) * rows / 8 in order, so
    // the y and z neighbour rows a block reads were read moments before by
    // blocks of the same XCD and come from its L2
    int rbeg = (int)blockIdx.x * rpb, rstep = nblk * rpb, rend = rows;
    if (nblk >= kXcds && nblk % kXcds == 0) {
        const int k = (int)blockIdx.x % kXcds, t = (int)blockIdx.x / kXcds;
        rbeg = (int)((int64_t)rows * k / kXcds) + t * rpb;
        rstep = nblk / kXcds * rpb;
        rend = (int)((int64_t)rows * (k + 1) / kXcds);
    }
    for (int r = rbeg + tsub; tsub < rpb && r < rend; r += rstep) {
        const int iy = r % ny;
        const int iz = r / ny;
        const int64_t row = (int64_t)r * nx;
        const int64_t dyp = iy + 1 < ny ? sy : -(int64_t)(ny - 1) * sy;
        const int64_t dym = iy > 0 ? -sy : (int64_t)(ny - 1) * sy;
        const int64_t dzp = iz + 1 < nz ? sz : -(int64_t)(nz - 1) * sz;
        const int64_t dzm = iz > 0 ? -sz : (int64_t)(nz - 1) * sz;
        for (int ix = tx0; ix < nx; ix += xstep) {
            const int64_t o = row + ix;
            const int64_t oxm = ix > 0 ? o - 1 : row + nx - 1;
            const int64_t oxp = ix + V < nx ? o + V : row;
            if (V == 1) {
                // all eight loads are issued before the stores
                const double p0 = phi[o], q = rho[o];
                const double vx = 0.5 * ((phi[oxm] - p0) + (p0 - phi[oxp]));
                const double vy = 0.5 * ((phi[o + dym] - p0) + (p0 - phi[o + dyp]));
                const double vz = 0.5 * ((phi[o + dzm] - p0) + (p0 - phi[o + dzp]));
                efield_store(ex + o, vx);
                efield_store(ey + o, vy);
                efield_store(ez + o, vz);
                we += q * p0;
                we2 += vx * vx + vy * vy + vz * vz;
                qn += q;
            } else {
                // o is even, so every pair is 16-byte aligned
                const gmem_t<d2v> *p2 = (const gmem_t<d2v> *)phi;
                const d2v c = p2[o >> 1], q = ((const gmem_t<d2v> *)rho)[o >> 1];
                const d2v ym = p2[(o + dym) >> 1], yp = p2[(o + dyp) >> 1];
                const d2v zm = p2[(o + dzm) >> 1], zp = p2[(o + dzp) >> 1];
                const double xm = phi[oxm], xp = phi[oxp];
                d2v vx, vy, vz;
                vx.x = 0.5 * ((xm - c.x) + (c.x - c.y));
                vx.y = 0.5 * ((c.x - c.y) + (c.y - xp));
                vy.x = 0.5 * ((ym.x - c.x) + (c.x - yp.x));
                vy.y = 0.5 * ((ym.y - c.y) + (c.y - yp.y));
                vz.x = 0.5 * ((zm.x - c.x) + (c.x - zp.x));
                vz.y = 0.5 * ((zm.y - c.y) + (c.y - zp.y));
                efield_store((gmem_t<d2v> *)ex + (o >> 1), vx);
                efield_store((gmem_t<d2v> *)ey + (o >> 1), vy);
                efield_store((gmem_t<d2v> *)ez + (o >> 1), vz);
                we += q.x * c.x;
                we += q.y * c.y;
                we2 += vx.x * vx.x + vy.x * vy.x + vz.x * vz.x;
                we2 += vx.y * vx.y + vy.y * vy.y + vz.y * vz.y;
                qn += q.x;
                qn += q.y;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_efield_b(const InstDev *__restrict__ insts)
{
    __shared__ double red[3][4];
    const InstDev *I = insts + blockIdx.y;
    const int nx = I->n[0];
    if ((int)blockIdx.x >= efield_blocks_of(nx, I->n[1], I->n[2])) return;
    double we = 0.0, we2 = 0.0, qn = 0.0;
    if (efield_vec(nx) == 2)
        efield_rows<2>(I, we, we2, qn);
    else
        efield_rows<1>(I, we, we2, qn);
    we = wave_sum(we);
    we2 = wave_sum(we2);
    qn = wave_sum(qn);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        red[0][w] = we;
        red[1][w] = we2;
        red[2][w] = qn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            a += red[0][i];
            b += red[1][i];
            c += red[2][i];
        }
        I->e_partial[3 * blockIdx.x + 0] = a;
        I->e_partial[3 * blockIdx.x + 1] = b;
        I->e_partial[3 * blockIdx.x + 2] = c;
    }
}

// threads of the blocks that reduce per-item and per-block partial sums
constexpr int kSumBlock = 1024;

// Sum of partial[i*stride + k] over i < count by one block, in an order that
// depends on count alone.  Each thread keeps kSumLanes running sums (entries
// i, i + blockDim, ... of one trip go to sums 0, 1, ...), so that many loads
// are in flight; a single running sum made the reduction over the 32768 work
// items of a 256^3 step wait for one load at a time.
__device__ double block_ordered_sum(const double *__restrict__ partial, int count, int stride,
                                    int k, double *red)
{
    constexpr int kSumLanes = 8;
    const int nb = (int)blockDim.x;
    double acc[kSumLanes];
#pragma unroll
    for (int a = 0; a < kSumLanes; ++a) acc[a] = 0.0;
    int i = threadIdx.x;
    for (; i + (kSumLanes - 1) * nb < count; i += kSumLanes * nb) {
#pragma unroll
        for (int a = 0; a < kSumLanes; ++a)
            acc[a] += gload(partial, (int64_t)(i + a * nb) * stride + k);
    }
#pragma unroll
    for (int a = 0; a < kSumLanes; ++a)
        if (i + a * nb < count) acc[a] += gload(partial, (int64_t)(i + a * nb) * stride + k);
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < kSumLanes; ++a) s += acc[a];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// blockIdx.x = instance: diag[kDiagWe + k] := sum of the instance's e_partial rows
__global__ void __launch_bounds__(kSumBlock) k_reduce_e_b(const InstDev *__restrict__ insts)
{
    __shared__ double red[kSumBlock];
    const InstDev *I = insts + blockIdx.x;
    const int count = efield_blocks_of(I->n[0], I->n[1], I->n[2]);
    for (int k = 0; k < 3; ++k) {
        const double v = block_ordered_sum(I->e_partial, count, 3, k, red);
        if (threadIdx.x == 0) I->diag[kDiagWe + k] = v;
    }
}

// Kinetic-energy slot k of an instance (k < kMaxSpecies: before the push, else
// after; species k % kMaxSpecies) summed over its items: row k of
// ke_partial[kKeSlots][nitems], contiguous, so that one block streams it.
// Slots of absent species are zero (their rows are not written).
__device__ double ke_slot_sum(const InstDev *I, int k, double *red)
{
    if (k % kMaxSpecies >= I->nspecies) return 0.0;
    return block_ordered_sum(I->ke_partial + (int64_t)k * I->nitems, I->nitems, 1, 0, red);
}

// blockIdx.y = instance, blockIdx.x = slot k: diag[kDiagKeOld + k] := sum
// over the instance's items of its kinetic-energy slot k
__global__ void __launch_bounds__(kSumBlock) k_reduce_ke_b(const InstDev *__restrict__ insts)
{
    __shared__ double red[kSumBlock];
    const InstDev *I = insts + blockIdx.y;
    const int k = blockIdx.x;
    const double v = ke_slot_sum(I, k, red);
    if (threadIdx.x == 0) I->diag[kDiagKeOld + k] = v;
}

__global__ void k_rho_from_acc_b(const InstDev *__restrict__ insts, const int *__restrict__ parity)
{
    const InstDev *I = insts + blockIdx.y;
    const unsigned long long *__restrict__ acc = I->rho_acc[parity[blockIdx.y]];
    double *__restrict__ rho = I->rho;
    const int64_t n = (int64_t)I->n[0] * I->n[1] * I->n[2];
    const double scale = I->fix_inv;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        rho[i] = (double)(long long)acc[i] * scale;
}

__global__ void k_prepare_next_b(const InstDev *__restrict__ insts, const int *__restrict__ parity,
                                 unsigned int *__restrict__ queue)
{
    // work counter of the queued particle launch that follows
    if (queue && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *queue = 0u;
    const InstDev *I = insts + blockIdx.y;
    const int nb = parity[blockIdx.y] ^ 1;
    const int64_t n = (int64_t)I->n[0] * I->n[1] * I->n[2];
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long *__restrict__ acc = I->rho_acc[nb];
    for (int64_t i = t0; i < n; i += stride) acc[i] = 0ull;
    for (int s = 0; s < I->nspecies; ++s) {
        unsigned long long *__restrict__ w = I->sp[s].count[nb];
        for (int64_t t = t0; t < I->ntiles; t += stride) w[t] = 0ull;
    }
    if (t0 < kFlagWords) I->flags[t0] = 0ull;
}

// Start of a step: rho := the current charge grid * 2^-charge_bits, and the
// zeroing of k_prepare_next_b (next charge grid and bin words, flags, queue
// counter) in the same sweep.
__global__ void k_begin_step_b(const InstDev *__restrict__ insts, const int *__restrict__ parity,
                               unsigned int *__restrict__ queue)
{
    if (queue && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *queue = 0u;
    const InstDev *I = insts + blockIdx.y;
    const int par = parity[blockIdx.y];
    const int64_t n = (int64_t)I->n[0] * I->n[1] * I->n[2];
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const unsigned long long *__restrict__ acc = I->rho_acc[par];
    unsigned long long *__restrict__ next = I->rho_acc[par ^ 1];
    double *__restrict__ rho = I->rho;
    const double scale = I->fix_inv;
    for (int64_t i = t0; i < n; i += stride) {
        rho[i] = (double)(long long)acc[i] * scale;
        next[i] = 0ull;
    }
    for (int s = 0; s < I->nspecies; ++s) {
        unsigned long long *__restrict__ w = I->sp[s].count[par ^ 1];
        for (int64_t t = t0; t < I->ntiles; t += stride) w[t] = 0ull;
    }
    if (t0 < kFlagWords) I->flags[t0] = 0ull;
}

// End of a step: blocks x < kKeSlots reduce the kinetic-energy slots (as
// k_reduce_ke_b), block kKeSlots the field-energy partials of k_efield_b (as
// k_reduce_e_b); unless a bin overflowed, all of them then scan a share of
// the step's new bins (buffer parity ^ 1): listed bins go to tight_list (their
// number to kFlagListed) and a tight one raises kFlagTight (the host then
// moves them before the next step).  When out is not null the last block to
// finish completes the instance's step in a group: out[i*kFlagSlots + k] :=
// flags[k] of instance i, and parity[i] flips unless a bin overflowed (the host
// then redoes that step in list mode from the unchanged buffer).
__global__ void __launch_bounds__(kSumBlock) k_end_step_b(const InstDev *__restrict__ insts,
                                                    int *__restrict__ parity,
                                                    unsigned long long *__restrict__ out)
{
    __shared__ double red[kSumBlock];
    const InstDev *I = insts + blockIdx.y;
    if ((int)blockIdx.x < kKeSlots) {
        const int k = blockIdx.x;
        const double v = ke_slot_sum(I, k, red);
        if (threadIdx.x == 0) I->diag[kDiagKeOld + k] = v;
    } else {
        const int count = efield_blocks_of(I->n[0], I->n[1], I->n[2]);
        for (int k = 0; k < 3; ++k) {
            const double v = block_ordered_sum(I->e_partial, count, 3, k, red);
            if (threadIdx.x == 0) I->diag[kDiagWe + k] = v;
        }
    }
    unsigned long long *flags = I->flags;
    if (flags[kFlagOverflow] == 0ull) {
        const int nb = parity[blockIdx.y] ^ 1;
        const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
        const int stride = gridDim.x * blockDim.x;
        for (int s = 0; s < I->nspecies; ++s) {
            const SpecDev &S = I->sp[s];
            for (int t = t0; t < I->ntiles; t += stride) {
                const unsigned long long c = S.count[nb][t];
                const int cap = S.bin_cap[t], used = bin_front(c) + bin_back(c);
                if (!bin_listed(cap, used, I->tight_frac)) continue;
                const unsigned long long k = atomicAdd(&flags[kFlagListed], 1ull);
                if (k < (unsigned long long)kTightListCap) I->tight_list[k] = tight_entry(s, t, used);
                if (bin_tight(cap, used, I->tight_frac)) atomicOr(&flags[kFlagTight], 1ull);
                __threadfence();  // visible before this block counts itself done
            }
        }
    }
    if (!out) return;
    // the flags are device-scope atomics (a thread that listed a bin fenced
    // its writes before the barrier), so the counter needs no fence of its
    // own: device-scope fences write back the L2 and cost microseconds per
    // block on this multi-XCD device
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0)
        last = atomicAdd(&flags[kFlagEndBlocks], 1ull) == (unsigned long long)gridDim.x - 1ull;
    __syncthreads();
    if (last && threadIdx.x == 0) {
        const int i = blockIdx.y;
        unsigned long long f[kFlagSlots];
        for (int k = 0; k < kFlagSlots; ++k) {
            f[k] = __hip_atomic_load(&flags[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[i * kFlagSlots + k] = f[k];
        }
        if (f[kFlagOverflow] == 0ull) parity[i] ^= 1;
    }
}

__global__ void k_load_uniform(Pcg32Dev rng0, uint64_t offset, int64_t np, int nx, int ny,
                               int nz, double vth, double *__restrict__ x, double *__restrict__ y,
                               double *__restrict__ z, double *__restrict__ vx,
                               double *__restrict__ vy, double *__restrict__ vz,
                               int64_t *__restrict__ pid, int64_t pid0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Pcg32Dev r = rng0;
    pcg_advance(r, offset + 7u * (uint64_t)i);
    // position_random_uniform_distribution.f90:41 with lower = 0, upper = n
    x[i] = pcg_double(r) * (double)nx + 0.0;
    y[i] = pcg_double(r) * (double)ny + 0.0;
    z[i] = pcg_double(r) * (double)nz + 0.0;
    maxwell3(r, vth, vx[i], vy[i], vz[i]);
    if (pid) pid[i] = pid0 + i;
}

__global__ void k_load_hump(Pcg32Dev rng0, uint64_t offset, uint64_t budget, int64_t np, int nx,
                            int ny, int nz, double amp, double sigma, double vth,
                            double *__restrict__ x, double *__restrict__ y,
                            double *__restrict__ z, double *__restrict__ vx,
                            double *__restrict__ vy, double *__restrict__ vz,
                            int64_t *__restrict__ pid, int64_t pid0,
                            unsigned long long *__restrict__ fail)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Pcg32Dev r = rng0;
    pcg_advance(r, offset + (uint64_t)i * budget);
    const double cx = 0.5 * nx, cy = 0.5 * ny, cz = 0.5 * nz;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    const uint64_t maxc = (budget - 4u) / 4u;
    bool ok = false;
    for (uint64_t c = 0; c < maxc; ++c) {
        const double px = pcg_double(r) * (double)nx;
        const double py = pcg_double(r) * (double)ny;
        const double pz = pcg_double(r) * (double)nz;
        const double u = pcg_double(r);
        const double dx = px - cx, dy = py - cy, dz = pz - cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double f = (1.0 + amp * exp(-r2 * inv2s2)) / (1.0 + amp);
        if (u < f) {
            x[i] = px;
            y[i] = py;
            z[i] = pz;
            ok = true;
            break;
        }
    }
    if (!ok) {
        atomicAdd(fail, 1ull);
        x[i] = 0.5 * nx;
        y[i] = 0.5 * ny;
        z[i] = 0.5 * nz;
    }
    maxwell3(r, vth, vx[i], vy[i], vz[i]);
    if (pid) pid[i] = pid0 + i;
}

}  // namespace

static unsigned grid_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, int64_t n, double v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

static unsigned grid_cap(int64_t n, int64_t cap)
{
    int64_t b = (n + 255) / 256;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// Blocks per instance of the grid-stride grid kernels: enough for n
// elements, at most 4096, and within a budget of about 16K blocks per launch
// shared by the instances (small instances need few).
static unsigned grid_blocks(int64_t n, int ninst)
{
    const int64_t budget = std::max<int64_t>(32, 16384 / std::max(ninst, 1));
    return grid_cap(n, std::min<int64_t>(4096, budget));
}

hipError_t launch_rho_from_acc_b(const InstDev *insts, const int *parity, int ninst,
                                 int64_t max_ng, hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rho_from_acc_b, dim3(grid_blocks(max_ng, ninst), ninst), dim3(256), 0,
                       st, insts, parity);
    return hipGetLastError();
}

hipError_t launch_efield_b(const InstDev *insts, int ninst, int max_blocks, bool reduce,
                           hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    const int blocks = (int)round_up_xcds(max_blocks > 0 ? max_blocks : 1);
    hipLaunchKernelGGL(k_efield_b, dim3(blocks, ninst), dim3(256), 0, st, insts);
    if (reduce) hipLaunchKernelGGL(k_reduce_e_b, dim3(ninst), dim3(kSumBlock), 0, st, insts);
    return hipGetLastError();
}

hipError_t launch_begin_step_b(const InstDev *insts, const int *parity, int ninst,
                               int64_t max_ng, int max_ntiles, unsigned int *queue,
                               hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    const int64_t n = max_ng > max_ntiles ? max_ng : max_ntiles;
    hipLaunchKernelGGL(k_begin_step_b, dim3(grid_blocks(n, ninst), ninst), dim3(256), 0, st,
                       insts, parity, queue);
    return hipGetLastError();
}

hipError_t launch_end_step_b(const InstDev *insts, int *parity, int ninst,
                             unsigned long long *out, hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_end_step_b, dim3(kKeSlots + 1, ninst), dim3(kSumBlock), 0, st, insts,
                       parity, out);
    return hipGetLastError();
}

hipError_t launch_prepare_next_b(const InstDev *insts, const int *parity, int ninst,
                                 int64_t max_ng, int max_ntiles, unsigned int *queue,
                                 hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    const int64_t n = max_ng > max_ntiles ? max_ng : max_ntiles;
    hipLaunchKernelGGL(k_prepare_next_b, dim3(grid_blocks(n, ninst), ninst), dim3(256), 0, st,
                       insts, parity, queue);
    return hipGetLastError();
}

hipError_t launch_reduce_ke_b(const InstDev *insts, int ninst, hipStream_t st)
{
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_ke_b, dim3(kKeSlots, ninst), dim3(kSumBlock), 0, st, insts);
    return hipGetLastError();
}

hipError_t launch_load_uniform(Pcg32Dev rng0, uint64_t offset, int64_t np, int nx, int ny,
                               int nz, double vth, double *x, double *y, double *z, double *vx,
                               double *vy, double *vz, int64_t *pid, int64_t pid0,
                               hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    const int64_t blocks = (np + 255) / 256;
    hipLaunchKernelGGL(k_load_uniform, dim3((unsigned)blocks), dim3(256), 0, st, rng0, offset,
                       np, nx, ny, nz, vth, x, y, z, vx, vy, vz, pid, pid0);
    return hipGetLastError();
}

hipError_t launch_load_hump(Pcg32Dev rng0, uint64_t offset, uint64_t budget, int64_t np, int nx,
                            int ny, int nz, double amp, double sigma, double vth, double *x,
                            double *y, double *z, double *vx, double *vy, double *vz,
                            int64_t *pid, int64_t pid0, unsigned long long *fail,
                            hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    const int64_t blocks = (np + 255) / 256;
    hipLaunchKernelGGL(k_load_hump, dim3((unsigned)blocks), dim3(256), 0, st, rng0, offset,
                       budget, np, nx, ny, nz, amp, sigma, vth, x, y, z, vx, vy, vz, pid, pid0,
                       fail);
    return hipGetLastError();
}

}  // namespace ohpic
