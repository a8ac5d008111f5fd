// ohpic_fft.hpp -- axis-pass FFT launchers (FFTW R2HC/HC2R semantics) and
// the per-length plan (radix factors + twiddle table in device memory).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ohpic {

constexpr int kMaxRadix = 13;   // largest prime factor an axis length may have
constexpr int kMaxFactors = 24;
constexpr int kFftBlock = 256;     // threads per FFT workgroup
constexpr int kFftLdsCplx = 2048;  // complex values per workgroup (32 KiB LDS)
constexpr int kMaxFftLen = 2048;

enum FftKind : int { kFftR2HC = 0, kFftHC2R = 1, kFftPoisson = 2 };

struct FftAxisDev {
    int n;
    int nfac;
    int fac[kMaxFactors];
    const double2 *tw;  // tw[k] = exp(-2 pi i k / n), k in [0, n)
};

// Factorises n (radix 4 first, then 2, 3, 5, ...); false if a prime factor
// exceeds kMaxRadix or n is out of range.
bool fft_factorize(int n, int *fac, int *nfac);

// Host-side twiddle table exp(-2 pi i k/n) in long double, rounded to double.
void fft_twiddles(int n, double2 *tw);

int fft_group_lines(int n);
size_t fft_lds_bytes(int n, int G);
int64_t fft_groups(int axis, int G, int nx, int ny, int nz);

// One pass over every line of `axis`.  kind R2HC/HC2R transform src -> dst
// (in place allowed); kind Poisson (axis 2 only) runs forward, division by
// lamx[kx]+lamy[ky]+lamz[kz] (DC -> 0) and inverse along z.  When
// src_is_fixed the source is an int64 fixed-point grid multiplied by
// src_scale on load.  Every output is multiplied by out_scale.
hipError_t fft_axis_pass(int kind, int axis, const void *src, bool src_is_fixed,
                         double src_scale, double *dst, double out_scale, int nx, int ny,
                         int nz, const FftAxisDev &ax, const double *lamx, const double *lamy,
                         const double *lamz, hipStream_t st);

}  // namespace ohpic
