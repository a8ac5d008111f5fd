// ohpic_particle_common.hpp -- device helpers shared by the FP64 and the
// mixed-precision particle kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ohpic_internal.hpp"

namespace ohpic {
namespace {

// Minimum waves per SIMD the fused kernel is compiled for (register budget).
#ifndef OHPIC_PUSH_MIN_WAVES
#define OHPIC_PUSH_MIN_WAVES 1
#endif

constexpr int kWaves = kBlock / 64;
constexpr int kLeaveCap = 32;  // per-wave buffer of particles that left the tile

template <int WO>
__device__ __forceinline__ void cic_w(double r, double &w0, double &w1)
{
    // WO 0: lower node gets r (the reference's order); WO 1: lower gets 1-r
    if (WO == 0) {
        w0 = r;
        w1 = 1.0 - r;
    } else {
        w0 = 1.0 - r;
        w1 = r;
    }
}

// periodic_particle_boundary.f90:46-52: the two comparisons in this order
__device__ __forceinline__ double apply_periodic(double p, int n)
{
    const double dn = (double)n;
    if (p < 0) p = dn + p;
    if (p >= dn) p = p - dn;
    return p;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWaves; ++i) s += red[i];
    __syncthreads();
    return s;
}

}  // namespace
}  // namespace ohpic
