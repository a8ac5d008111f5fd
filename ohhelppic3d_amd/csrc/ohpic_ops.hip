// ohpic_ops.hip -- the reference's per-particle operators as batched gfx950
// kernels over caller-owned host arrays.
//
// These back the per-particle bindings of the Fortran types (move, interp,
// scatter, apply), which the reference calls once per particle
// (ohhelppic3d.f90:317,347,349,353,355).  Each call uploads the batch, runs
// one kernel and downloads the result; the whole-simulation path (ohpic_step)
// never goes through them.
//   k_boris_batch   particle_mover_boris.f90:21-64 (eb gathered per particle)
//   k_interp_batch  linear_interpolator.f90:22-69 over a t_OhField-shaped
//                   array values(nelem, lb:ub) (Fortran order)
//   k_scatter_batch linear_scatter.f90:22-55 into such an array
//   k_wrap_batch    periodic_particle_boundary.f90:42-53, applied k times
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/ohpic.h"
#include "ohpic_internal.hpp"
#include "ohpic_particle_common.hpp"

namespace ohpic {

namespace {

struct FieldBox {
    int nelem;
    int lb[3], ub[3];
    int64_t sx, sy;  // extents of axes 0 and 1
};

__device__ __forceinline__ int64_t fidx(const FieldBox &f, int c, int ix, int iy, int iz)
{
    return c + (int64_t)f.nelem *
                   ((ix - f.lb[0]) + f.sx * ((iy - f.lb[1]) + f.sy * (int64_t)(iz - f.lb[2])));
}

__device__ __forceinline__ bool stencil_inside(const FieldBox &f, const int i0[3])
{
    for (int a = 0; a < 3; ++a)
        if (i0[a] < f.lb[a] || i0[a] + 1 > f.ub[a]) return false;
    return true;
}

__global__ void k_boris_batch(int64_t n, double *__restrict__ x, double *__restrict__ y,
                              double *__restrict__ z, double *__restrict__ vx,
                              double *__restrict__ vy, double *__restrict__ vz,
                              const double *__restrict__ eb, double qm, double dt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *e = eb + 6 * i;
    const double dt2 = dt * 0.5;
    const double t0 = qm * e[3] * dt2, t1 = qm * e[4] * dt2, t2 = qm * e[5] * dt2;
    const double tt = t0 * t0 + t1 * t1 + t2 * t2;
    const double s0 = 2 * t0 / (1 + tt), s1 = 2 * t1 / (1 + tt), s2 = 2 * t2 / (1 + tt);
    const double um0 = vx[i] + qm * e[0] * dt2;
    const double um1 = vy[i] + qm * e[1] * dt2;
    const double um2 = vz[i] + qm * e[2] * dt2;
    const double ua0 = um0 + (um1 * t2 - um2 * t1);
    const double ua1 = um1 + (um2 * t0 - um0 * t2);
    const double ua2 = um2 + (um0 * t1 - um1 * t0);
    const double up0 = um0 + (ua1 * s2 - ua2 * s1);
    const double up1 = um1 + (ua2 * s0 - ua0 * s2);
    const double up2 = um2 + (ua0 * s1 - ua1 * s0);
    vx[i] = up0 + qm * e[0] * dt2;
    vy[i] = up1 + qm * e[1] * dt2;
    vz[i] = up2 + qm * e[2] * dt2;
    x[i] = x[i] + vx[i] * dt;
    y[i] = y[i] + vy[i] * dt;
    z[i] = z[i] + vz[i] * dt;
}

template <int WO>
__global__ void k_interp_batch(int64_t n, const double *__restrict__ pos,
                               const double *__restrict__ values, FieldBox f,
                               double *__restrict__ out, unsigned long long *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int i0[3];
    double w[3][2];
    for (int a = 0; a < 3; ++a) {
        const double p = pos[3 * i + a];
        i0[a] = (int)p;
        cic_w<WO>(p - (double)i0[a], w[a][0], w[a][1]);
    }
    double *o = out + (int64_t)f.nelem * i;
    for (int c = 0; c < f.nelem; ++c) o[c] = 0.0;
    if (!stencil_inside(f, i0)) {
        atomicAdd(bad, 1ull);
        return;
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 2; ++k) {
                const double wt = w[0][a] * w[1][b] * w[2][k];
                for (int c = 0; c < f.nelem; ++c)
                    o[c] += values[fidx(f, c, i0[0] + a, i0[1] + b, i0[2] + k)] * wt;
            }
}

template <int WO>
__global__ void k_scatter_batch(int64_t n, const double *__restrict__ pos,
                                double *__restrict__ values, FieldBox f,
                                const double *__restrict__ amount,
                                unsigned long long *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int i0[3];
    double w[3][2];
    for (int a = 0; a < 3; ++a) {
        const double p = pos[3 * i + a];
        i0[a] = (int)p;
        cic_w<WO>(p - (double)i0[a], w[a][0], w[a][1]);
    }
    if (!stencil_inside(f, i0)) {
        atomicAdd(bad, 1ull);
        return;
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 2; ++k) {
                const double wt = w[0][a] * w[1][b] * w[2][k];
                for (int c = 0; c < f.nelem; ++c)
                    atomicAdd(&values[fidx(f, c, i0[0] + a, i0[1] + b, i0[2] + k)],
                              amount[c] * wt);
            }
}

__global__ void k_wrap_batch(int64_t n, double *__restrict__ p, int nn, int applications)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = p[i];
    for (int k = 0; k < applications; ++k) v = apply_periodic(v, nn);
    p[i] = v;
}

}  // namespace

}  // namespace ohpic

using namespace ohpic;

namespace {

// Scoped device buffers + stream for one batch call.
struct Batch {
    std::vector<void *> ptrs;
    hipStream_t st = nullptr;
    hipError_t err = hipSuccess;
    explicit Batch(int device)
    {
        err = hipSetDevice(device);
        if (err == hipSuccess) err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    }
    ~Batch()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void *p : ptrs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T>
    T *upload(const T *h, int64_t n)
    {
        if (err != hipSuccess) return nullptr;
        T *d = nullptr;
        err = hipMalloc((void **)&d, (size_t)(n > 0 ? n : 1) * sizeof(T));
        if (err != hipSuccess) return nullptr;
        ptrs.push_back(d);
        if (h && n > 0) err = hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st);
        return d;
    }
    template <typename T>
    void download(T *h, const T *d, int64_t n)
    {
        if (err == hipSuccess && n > 0)
            err = hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, st);
    }
    int finish()
    {
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        return err == hipSuccess ? OHPIC_OK : OHPIC_EHIP;
    }
};

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

bool make_box(int nelem, const int32_t lb[3], const int32_t ub[3], FieldBox &f, int64_t &size)
{
    if (nelem < 1 || !lb || !ub) return false;
    f.nelem = nelem;
    size = nelem;
    for (int a = 0; a < 3; ++a) {
        if (ub[a] < lb[a]) return false;
        f.lb[a] = lb[a];
        f.ub[a] = ub[a];
        size *= (int64_t)(ub[a] - lb[a] + 1);
    }
    f.sx = ub[0] - lb[0] + 1;
    f.sy = ub[1] - lb[1] + 1;
    return true;
}

}  // namespace

extern "C" {

int ohpic_boris_batch(int device, int64_t n, double *x, double *y, double *z, double *vx,
                      double *vy, double *vz, const double *eb, double qm, double dt)
{
    if (n < 0 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !eb))) {
        set_global_error("boris_batch: bad arguments");
        return OHPIC_EINVAL;
    }
    if (n == 0) return OHPIC_OK;
    Batch b(device);
    double *d[6];
    double *h[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < 6; ++k) d[k] = b.upload(h[k], n);
    const double *de = b.upload(eb, 6 * n);
    if (b.err == hipSuccess)
        hipLaunchKernelGGL(k_boris_batch, dim3(blocks_for(n)), dim3(256), 0, b.st, n, d[0],
                           d[1], d[2], d[3], d[4], d[5], de, qm, dt);
    for (int k = 0; k < 6; ++k) b.download(h[k], d[k], n);
    const int rc = b.finish();
    if (rc) set_global_error(hipGetErrorString(b.err));
    return rc;
}

int ohpic_interp_batch(int device, int64_t n, const double *pos, const double *values, int nelem,
                       const int32_t lb[3], const int32_t ub[3], int weight_order, double *out)
{
    FieldBox f;
    int64_t size = 0;
    if (n < 0 || !make_box(nelem, lb, ub, f, size) || (n > 0 && (!pos || !values || !out)) ||
        (weight_order != 0 && weight_order != 1)) {
        set_global_error("interp_batch: bad arguments");
        return OHPIC_EINVAL;
    }
    if (n == 0) return OHPIC_OK;
    Batch b(device);
    const double *dp = b.upload(pos, 3 * n);
    const double *dv = b.upload(values, size);
    double *dout = b.upload<double>(nullptr, (int64_t)nelem * n);
    const unsigned long long zero = 0;
    unsigned long long *dbad = b.upload(&zero, 1);
    if (b.err == hipSuccess) {
        if (weight_order == 0)
            hipLaunchKernelGGL(k_interp_batch<0>, dim3(blocks_for(n)), dim3(256), 0, b.st, n, dp,
                               dv, f, dout, dbad);
        else
            hipLaunchKernelGGL(k_interp_batch<1>, dim3(blocks_for(n)), dim3(256), 0, b.st, n, dp,
                               dv, f, dout, dbad);
    }
    unsigned long long bad = 0;
    b.download(out, dout, (int64_t)nelem * n);
    b.download(&bad, dbad, 1);
    int rc = b.finish();
    if (rc) {
        set_global_error(hipGetErrorString(b.err));
        return rc;
    }
    if (bad) {
        set_global_error("interp_batch: a particle's stencil lies outside lb:ub");
        return OHPIC_EINVAL;
    }
    return OHPIC_OK;
}

int ohpic_scatter_batch(int device, int64_t n, const double *pos, double *values, int nelem,
                        const int32_t lb[3], const int32_t ub[3], const double *amount,
                        int weight_order)
{
    FieldBox f;
    int64_t size = 0;
    if (n < 0 || !make_box(nelem, lb, ub, f, size) || (n > 0 && (!pos || !values || !amount)) ||
        (weight_order != 0 && weight_order != 1)) {
        set_global_error("scatter_batch: bad arguments");
        return OHPIC_EINVAL;
    }
    if (n == 0) return OHPIC_OK;
    Batch b(device);
    const double *dp = b.upload(pos, 3 * n);
    double *dv = b.upload(values, size);
    const double *da = b.upload(amount, nelem);
    const unsigned long long zero = 0;
    unsigned long long *dbad = b.upload(&zero, 1);
    if (b.err == hipSuccess) {
        if (weight_order == 0)
            hipLaunchKernelGGL(k_scatter_batch<0>, dim3(blocks_for(n)), dim3(256), 0, b.st, n,
                               dp, dv, f, da, dbad);
        else
            hipLaunchKernelGGL(k_scatter_batch<1>, dim3(blocks_for(n)), dim3(256), 0, b.st, n,
                               dp, dv, f, da, dbad);
    }
    unsigned long long bad = 0;
    b.download(values, dv, size);
    b.download(&bad, dbad, 1);
    int rc = b.finish();
    if (rc) {
        set_global_error(hipGetErrorString(b.err));
        return rc;
    }
    if (bad) {
        set_global_error("scatter_batch: a particle's stencil lies outside lb:ub");
        return OHPIC_EINVAL;
    }
    return OHPIC_OK;
}

int ohpic_wrap_batch(int device, int64_t n, double *p, int32_t nn, int applications)
{
    if (n < 0 || nn < 1 || applications < 0 || (n > 0 && !p)) {
        set_global_error("wrap_batch: bad arguments");
        return OHPIC_EINVAL;
    }
    if (n == 0) return OHPIC_OK;
    Batch b(device);
    double *dp = b.upload(p, n);
    if (b.err == hipSuccess)
        hipLaunchKernelGGL(k_wrap_batch, dim3(blocks_for(n)), dim3(256), 0, b.st, n, dp, nn,
                           applications);
    b.download(p, dp, n);
    const int rc = b.finish();
    if (rc) set_global_error(hipGetErrorString(b.err));
    return rc;
}

}  // extern "C"
