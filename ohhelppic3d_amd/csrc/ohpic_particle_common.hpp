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

// The value lane `src` holds, for a wave-uniform `src`: v_readlane through a
// scalar register (a shuffle would route all 64 lanes through the LDS unit,
// which the gather and deposit keep busy).
__device__ __forceinline__ int lane_bcast(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ __forceinline__ unsigned long long lane_bcast(unsigned long long v, int src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}

// Particle state streams through each step once: with OHPIC_NT set, its
// loads and stores carry the non-temporal hint (measured 3-7% slower on the
// 256^3 x 64 ppc step, so off by default).
#ifndef OHPIC_NT
#define OHPIC_NT 0
#endif

template <typename T>
__device__ __forceinline__ T sload(const T *p, int64_t i)
{
    if (OHPIC_NT) return __builtin_nontemporal_load(&as_global(p)[i]);
    return as_global(p)[i];
}

template <typename T>
__device__ __forceinline__ void sstore(T *p, int64_t i, T v)
{
    if (OHPIC_NT)
        __builtin_nontemporal_store(v, &as_global(p)[i]);
    else
        as_global(p)[i] = v;
}

// relaxed device-scope atomic add on global memory (as atomicAdd)
__device__ __forceinline__ unsigned long long gatomic_add(unsigned long long *p,
                                                          unsigned long long v)
{
    return __hip_atomic_fetch_add(as_global(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// counter[t] += 1 for every active lane with `valid`; returns the lane's
// value of the counter before its own increment.  Lanes of the wavefront
// that share t are served by one atomic (their leader adds the group size and
// each lane takes old + its rank in the group), for the first kAggRounds
// distinct values of t; the remaining lanes add individually.  Particle lists
// written by the step kernels keep each source tile's particles together, so
// a wavefront usually targets one to a few bins.
constexpr int kAggRounds = 4;

__device__ __forceinline__ unsigned long long wave_counter_add(unsigned long long *counter, int t,
                                                               bool valid)
{
    const int lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(valid);
    unsigned long long result = 0ull;
    for (int r = 0; r < kAggRounds && todo; ++r) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int tl = lane_bcast(t, leader);
        const unsigned long long m = __ballot(valid && t == tl) & todo;
        unsigned long long old = 0ull;
        if (lane == leader) old = gatomic_add(&counter[tl], (unsigned long long)__popcll(m));
        old = lane_bcast(old, leader);
        if ((m >> lane) & 1ull) result = old + (unsigned long long)__popcll(m & lt);
        todo &= ~m;
    }
    if ((todo >> lane) & 1ull) result = gatomic_add(&counter[t], 1ull);
    return result;
}

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
