// ohpic_fields.hip -- grid kernels (fills, charge conversion, node-centred E
// with energy diagnostics) and the seeded particle loaders for gfx950.
#include <hip/hip_runtime.h>

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

__global__ void k_rho_from_acc(const unsigned long long *__restrict__ acc, double *__restrict__ rho,
                               int64_t n, double scale)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        rho[i] = (double)(long long)acc[i] * scale;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// E_a(i) = 0.5*((phi(i-e_a) - phi(i)) + (phi(i) - phi(i+e_a))): the edge
// difference of poisson_field_solver.f90:139-144 averaged onto nodes as at
// :165-170, periodic.  Each block walks whole x rows (32-bit row indexing)
// and writes its partial sums of rho*phi, |E|^2 and rho to partial[3*block];
// k_reduce_partials adds them in block order.
__global__ void __launch_bounds__(256)
    k_efield(const double *__restrict__ phi, const double *__restrict__ rho,
             double *__restrict__ ex, double *__restrict__ ey, double *__restrict__ ez, int nx,
             int ny, int nz, double *__restrict__ partial)
{
    __shared__ double red[3][4];
    const int rows = ny * nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    double we = 0.0, we2 = 0.0, qn = 0.0;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int iy = r % ny;
        const int iz = r / ny;
        const int64_t row = (int64_t)r * nx;
        const int64_t dyp = iy + 1 < ny ? sy : -(int64_t)(ny - 1) * sy;
        const int64_t dym = iy > 0 ? -sy : (int64_t)(ny - 1) * sy;
        const int64_t dzp = iz + 1 < nz ? sz : -(int64_t)(nz - 1) * sz;
        const int64_t dzm = iz > 0 ? -sz : (int64_t)(nz - 1) * sz;
        for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
            const int64_t o = row + ix;
            const int64_t oxp = ix + 1 < nx ? o + 1 : row;
            const int64_t oxm = ix > 0 ? o - 1 : row + nx - 1;
            const double p0 = phi[o];
            const double vx = 0.5 * ((phi[oxm] - p0) + (p0 - phi[oxp]));
            const double vy = 0.5 * ((phi[o + dym] - p0) + (p0 - phi[o + dyp]));
            const double vz = 0.5 * ((phi[o + dzm] - p0) + (p0 - phi[o + dzp]));
            ex[o] = vx;
            ey[o] = vy;
            ez[o] = vz;
            const double q = rho[o];
            we += q * p0;
            we2 += vx * vx + vy * vy + vz * vz;
            qn += q;
        }
    }
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
        partial[3 * blockIdx.x + 0] = a;
        partial[3 * blockIdx.x + 1] = b;
        partial[3 * blockIdx.x + 2] = c;
    }
}

// out[k] (+)= sum over i of partial[i*stride + k], k < width, summed in index
// order by one block (deterministic).
__global__ void __launch_bounds__(256)
    k_reduce_partials(const double *__restrict__ partial, int count, int stride, int width,
                      double *__restrict__ out, int accumulate)
{
    __shared__ double red[256];
    for (int k = 0; k < width; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[(int64_t)i * stride + k];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = (accumulate ? out[k] : 0.0) + red[0];
        __syncthreads();
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

hipError_t launch_rho_from_acc(const unsigned long long *acc, double *rho, int64_t n,
                               double scale, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rho_from_acc, dim3(grid_for(n)), dim3(256), 0, st, acc, rho, n, scale);
    return hipGetLastError();
}

int efield_blocks(int ny, int nz)
{
    const int rows = ny * nz;
    return rows < kEfieldMaxBlocks ? rows : kEfieldMaxBlocks;
}

hipError_t launch_efield(const double *phi, const double *rho, double *ex, double *ey,
                         double *ez, int nx, int ny, int nz, double *partial, double *diag,
                         hipStream_t st)
{
    const int blocks = efield_blocks(ny, nz);
    hipLaunchKernelGGL(k_efield, dim3(blocks), dim3(256), 0, st, phi, rho, ex, ey, ez, nx, ny,
                       nz, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, partial, blocks, 3, 3,
                       diag + kDiagWe, 0);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double *partial, int count, int stride, int width,
                                  double *out, int accumulate, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, partial, count, stride,
                       width, out, accumulate);
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
