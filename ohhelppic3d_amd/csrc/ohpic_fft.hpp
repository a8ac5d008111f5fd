// ohpic_fft.hpp -- axis-pass FFT launchers (FFTW R2HC/HC2R semantics) and
// the per-length plan (radix factors + twiddle table in device memory).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ohpic {

constexpr int kMaxRadix = 13;   // largest prime factor an axis length may have
constexpr int kMaxFactors = 24;
#ifndef OHPIC_FFT_BLOCK
#define OHPIC_FFT_BLOCK 256
#endif
constexpr int kFftBlock = OHPIC_FFT_BLOCK;  // threads per FFT workgroup
// 1024 complex values (16 KiB) per workgroup: ten workgroups fit a CU's LDS
// and the passes run at 8 waves per SIMD (34-46 VGPRs); measured 28% faster
// per pass at 256^3 than 32 KiB at 4 waves.  Axis lengths up to 1024.
#ifndef OHPIC_FFT_LDS_CPLX
#define OHPIC_FFT_LDS_CPLX 1024
#endif
constexpr int kFftLdsCplx = OHPIC_FFT_LDS_CPLX;  // complex values per workgroup (16 B each)
constexpr int kMaxFftLen = kFftLdsCplx;  // one complex line must fit a workgroup's LDS

enum FftKind : int { kFftR2HC = 0, kFftHC2R = 1, kFftPoisson = 2 };

// q = t / d for 0 <= t < 2^16 and 1 <= d <= 2^16 as one high multiply by
// m = ceil(2^32 / d) (m = 0 encodes d = 1): with t = q d + r, t m / 2^32 =
// q + r/d + t e/(d 2^32) where m d = 2^32 + e, e < d, and t d < 2^32 keeps the
// last term below 1/d, so the floor is exact.
__host__ __device__ inline uint32_t fastdiv_magic(uint32_t d)
{
    return d <= 1 ? 0u : (uint32_t)((0x100000000ull + d - 1) / d);
}

struct FftAxisDev {
    int n;
    int nfac;
    int fac[kMaxFactors];
    const double2 *tw;  // tw[k] = exp(-2 pi i k / n), k in [0, n)
    // divisors of the LDS passes as fastdiv multipliers (fft_plan_divisors)
    uint32_t mg_n;                 // n
    uint32_t mg_G;                 // fft_group_lines(n)
    uint32_t mg_half;              // n / 2 + 1
    uint32_t mg_nl;                // fft_group_lines(n) / 2: line pairs per group
    uint32_t mg_tpl[kMaxFactors];  // n / fac[s]: butterflies per line in stage s
    uint32_t mg_Ls[kMaxFactors];   // fac[0] * ... * fac[s-1]
};

// Factorises n (radix 4 first, then 2, 3, 5, ...); false if a prime factor
// exceeds kMaxRadix or n is out of range.
bool fft_factorize(int n, int *fac, int *nfac);

// Fills the fastdiv multipliers of a factorised axis plan.
void fft_plan_divisors(FftAxisDev &ax);

// Host-side twiddle table exp(-2 pi i k/n) in long double, rounded to double.
void fft_twiddles(int n, double2 *tw);

// real lines per workgroup: G/2 complex lines of n fill at most kFftLdsCplx
// complex values (32 KiB); at least 2 lines, at most 64
__host__ __device__ inline int fft_group_lines(int n)
{
    int g = 2 * (kFftLdsCplx / (n > 0 ? n : 1));
    if (g > 64) g = 64;
    if (g < 2) g = 2;
    return g & ~1;
}

// workgroups of one axis pass (G real lines each)
__host__ __device__ inline int64_t fft_groups(int axis, int G, int nx, int ny, int nz)
{
    if (axis == 0) return ((int64_t)ny * nz + G - 1) / G;
    if (axis == 1) return (int64_t)((nx + G - 1) / G) * nz;
    return ((int64_t)nx * ny + G - 1) / G;
}

// One transform job: an nx*ny*nz array with its per-axis plans and (for the
// Poisson pass) the Green's-function denominators per axis.
struct FftJob {
    const double *src;     // source of the first Poisson pass (rho)
    double *data;          // array transformed in place (phi)
    int n[3];
    FftAxisDev ax[3];
    const double *lam[3];  // (2 sin(pi k/n))^2 per axis
    double inv_n;          // 1 / (nx ny nz)
};

enum FftPass : int {
    kPassR2hcXFromSrc = 0,  // R2HC along x, src -> data
    kPassR2hcX,             // R2HC along x in place
    kPassR2hcY,
    kPassR2hcZ,
    kPassPoissonZ,          // R2HC z, divide by lam sum (DC := 0), HC2R z
    kPassHc2rZ,
    kPassHc2rY,
    kPassHc2rXScaled        // HC2R along x, times inv_n
};

int fft_radix_set(const FftAxisDev &ax);

// Which jobs of a Poisson z pass have 256-point z lines (run by the 16 x 16
// register kernel; the others by the radix-4 LDS stages of k_fft_pass).
enum Z256Jobs : int { kZ256None = 0, kZ256All = 1, kZ256Some = 2 };

// The fused Poisson z pass (kPassPoissonZ) for every job of the table.
hipError_t fft_poisson_z(const FftJob *jobs, int njobs, int64_t max_groups, int rs, int z256,
                         hipStream_t st);

// One axis pass for every job of the device table `jobs`; max_groups is the
// largest group count over the jobs, rs the largest radix set.
hipError_t fft_pass(FftPass pass, const FftJob *jobs, int njobs, int64_t max_groups, int rs,
                    hipStream_t st);

}  // namespace ohpic
