// ohpic_pcg.hpp -- device PCG32 and the seeded Maxwellian draw, shared by
// the FP64 and mixed-precision loaders.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ohpic_kernels.hpp"

namespace ohpic {
namespace {

constexpr double kRefPi = 3.14159265358979;  // science_constants.f90:4

// ---------------------------------------------------------------- PCG32
// lib/pcg-c: pcg_setseq_64 step (pcg_variants.h:588-591), XSH-RR output
// (:156-159), advance (pcg-advance-64.c:44-59), double = ldexp(u32, -32)
// (pcg32_double.c:60-63).
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

__device__ __forceinline__ uint32_t pcg_next(Pcg32Dev &r)
{
    const uint64_t old = r.state;
    r.state = old * kPcgMult + r.inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

__device__ __forceinline__ double pcg_double(Pcg32Dev &r)
{
    return (double)pcg_next(r) * 2.3283064365386962890625e-10;  // 2^-32
}

__device__ __forceinline__ void pcg_advance(Pcg32Dev &r, uint64_t delta)
{
    uint64_t cm = kPcgMult, cp = r.inc, am = 1u, ap = 0u;
    while (delta > 0) {
        if (delta & 1u) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1u) * cp;
        cm *= cm;
        delta >>= 1u;
    }
    r.state = am * r.state + ap;
}

__device__ __forceinline__ void maxwell3(Pcg32Dev &r, double vth, double &v1, double &v2,
                                         double &v3)
{
    // four uniforms in the order of velocity_maxwell_distribution3d.f90:46-56;
    // std = vth and log(1-u)
    const double r1 = pcg_double(r), r2 = pcg_double(r), r3 = pcg_double(r),
                 r4 = pcg_double(r);
    const double a = vth * sqrt(-2.0 * log(1.0 - r1));
    const double b = vth * sqrt(-2.0 * log(1.0 - r3));
    v1 = a * cos(2.0 * kRefPi * r2);
    v2 = a * sin(2.0 * kRefPi * r2);
    v3 = b * sin(2.0 * kRefPi * r4);
}

}  // namespace
}  // namespace ohpic
