// ohpic_fft.hip -- separable real-to-real 3-D FFT (FFTW R2HC / HC2R
// semantics) and the periodic spectral Poisson solve for gfx950.
//
// Replaces the FFTW-MPI backend (mpi_fftw3_solver.f90:39-236) and
// poissonSolver3d_solve (poisson_solver3d.f90:134-207) on one device.
//
// Each launch transforms every line of one axis of an nx*ny*nz array (x
// fastest).  A workgroup loads G real lines of that axis with coalesced
// reads (G consecutive rows for x; G consecutive x positions for y and z),
// packs line pairs into complex lines z = a + i b in LDS, runs a mixed-radix
// Stockham FFT in LDS (radix 4/2/3/5 and a generic small-prime stage), and
// unpacks the two real spectra into FFTW's halfcomplex order
// r0, r1, ..., r_{n/2}, i_{(n+1)/2-1}, ..., i1.  The inverse runs the same
// steps backwards.  The z pass of the Poisson solve is fused: forward along
// z, divide by the modified wave number sum (DC := 0), inverse along z, in one
// LDS-resident visit of each z line.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ohpic_fft.hpp"

namespace ohpic {

namespace {

constexpr int kFftBlock = 256;

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// One Stockham radix-p stage over `nlines` complex lines of length n held in
// LDS.  Input layout per line: y_prev[(row + q*r)*Ls + f]; output
// y[row*L + f + Ls*u] = sum_q (w_L^{f q} y_prev[...]) w_p^{u q}.
template <int SIGN>
__device__ void fft_stage(const cplx *__restrict__ src, cplx *__restrict__ dst, int n,
                          int nlines, int p, int Ls, const double2 *__restrict__ tw)
{
    const int L = Ls * p;
    const int r = n / L;
    const int tasks_per_line = n / p;
    const int total = tasks_per_line * nlines;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int line = t / tasks_per_line;
        const int i = t - line * tasks_per_line;
        const int row = i / Ls;
        const int f = i - row * Ls;
        const cplx *s = src + line * n;
        cplx *d = dst + line * n;
        if (p == 4) {
            cplx a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cplx v = s[(row + q * r) * Ls + f];
                if (q > 0 && f > 0) {
                    double2 w = tw[f * q * r];
                    v = cmul(v, cplx{w.x, SIGN < 0 ? w.y : -w.y});
                }
                a[q] = v;
            }
            cplx t0 = {a[0].re + a[2].re, a[0].im + a[2].im};
            cplx t1 = {a[0].re - a[2].re, a[0].im - a[2].im};
            cplx t2 = {a[1].re + a[3].re, a[1].im + a[3].im};
            cplx t3 = {a[1].re - a[3].re, a[1].im - a[3].im};
            // w4 * t3 with w4 = SIGN * i
            cplx wt3 = {-SIGN * t3.im, SIGN * t3.re};
            const int ob = row * L + f;
            d[ob] = {t0.re + t2.re, t0.im + t2.im};
            d[ob + Ls] = {t1.re + wt3.re, t1.im + wt3.im};
            d[ob + 2 * Ls] = {t0.re - t2.re, t0.im - t2.im};
            d[ob + 3 * Ls] = {t1.re - wt3.re, t1.im - wt3.im};
        } else if (p == 2) {
            cplx a0 = s[row * Ls + f];
            cplx a1 = s[(row + r) * Ls + f];
            if (f > 0) {
                double2 w = tw[f * r];
                a1 = cmul(a1, cplx{w.x, SIGN < 0 ? w.y : -w.y});
            }
            const int ob = row * L + f;
            d[ob] = {a0.re + a1.re, a0.im + a1.im};
            d[ob + Ls] = {a0.re - a1.re, a0.im - a1.im};
        } else if (p == 3) {
            cplx a[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                cplx v = s[(row + q * r) * Ls + f];
                if (q > 0 && f > 0) {
                    double2 w = tw[f * q * r];
                    v = cmul(v, cplx{w.x, SIGN < 0 ? w.y : -w.y});
                }
                a[q] = v;
            }
            const double c = -0.5;
            const double sn = SIGN * 0.86602540378443864676372317075294;  // sin(2 pi/3)
            cplx tsum = {a[1].re + a[2].re, a[1].im + a[2].im};
            cplx tdif = {a[1].re - a[2].re, a[1].im - a[2].im};
            cplx m = {a[0].re + c * tsum.re, a[0].im + c * tsum.im};
            cplx j = {-sn * tdif.im, sn * tdif.re};  // i * sn * tdif
            const int ob = row * L + f;
            d[ob] = {a[0].re + tsum.re, a[0].im + tsum.im};
            d[ob + Ls] = {m.re + j.re, m.im + j.im};
            d[ob + 2 * Ls] = {m.re - j.re, m.im - j.im};
        } else {
            // generic small prime (p <= kMaxRadix)
            cplx a[kMaxRadix];
            for (int q = 0; q < p; ++q) {
                cplx v = s[(row + q * r) * Ls + f];
                if (q > 0 && f > 0) {
                    double2 w = tw[f * q * r];
                    v = cmul(v, cplx{w.x, SIGN < 0 ? w.y : -w.y});
                }
                a[q] = v;
            }
            const int step = n / p;
            const int ob = row * L + f;
            for (int u = 0; u < p; ++u) {
                cplx acc = a[0];
                for (int q = 1; q < p; ++q) {
                    double2 w = tw[((u * q) % p) * step];
                    cplx v = cmul(a[q], cplx{w.x, SIGN < 0 ? w.y : -w.y});
                    acc.re += v.re;
                    acc.im += v.im;
                }
                d[ob + Ls * u] = acc;
            }
        }
    }
}

// Runs all stages; returns the buffer holding the result.
template <int SIGN>
__device__ cplx *fft_lds(cplx *a, cplx *b, int n, int nlines, const FftAxisDev &ax)
{
    cplx *src = a, *dst = b;
    int Ls = 1;
    for (int s = 0; s < ax.nfac; ++s) {
        const int p = ax.fac[s];
        fft_stage<SIGN>(src, dst, n, nlines, p, Ls, ax.tw);
        __syncthreads();
        Ls *= p;
        cplx *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Line addressing for an axis pass.  Returns the base element of line g of
// group grp (or -1 when that line does not exist) and the element stride.
struct LineGeom {
    int64_t base;
    int ix, iy;  // coordinates used by the Poisson division (z pass)
};

template <int AXIS>
__device__ __forceinline__ bool line_of(int grp, int g, int G, int nx, int ny, int nz,
                                        LineGeom &lg)
{
    if (AXIS == 0) {
        const int64_t row = (int64_t)grp * G + g;
        if (row >= (int64_t)ny * nz) return false;
        lg.base = row * nx;
        lg.ix = 0;
        lg.iy = (int)(row % ny);
        return true;
    } else if (AXIS == 1) {
        const int gpp = (nx + G - 1) / G;
        const int iz = grp / gpp;
        const int ix = (grp - iz * gpp) * G + g;
        if (ix >= nx || iz >= nz) return false;
        lg.base = ix + (int64_t)nx * ny * iz;
        lg.ix = ix;
        lg.iy = 0;
        return true;
    } else {
        const int64_t l = (int64_t)grp * G + g;
        if (l >= (int64_t)nx * ny) return false;
        lg.base = l;
        lg.ix = (int)(l % nx);
        lg.iy = (int)(l / nx);
        return true;
    }
}

template <int AXIS>
__device__ __forceinline__ void elem_of(int e, int n, int G, int &g, int &j)
{
    if (AXIS == 0) {
        g = e / n;
        j = e - g * n;
    } else {
        j = e / G;
        g = e - j * G;
    }
}

// Load G real lines into LDS packed as complex (line pair -> re/im).
template <int AXIS, typename SrcT>
__device__ void load_lines(const SrcT *__restrict__ src, double src_scale, cplx *buf, int grp,
                           int G, int n, int64_t estride, int nx, int ny, int nz)
{
    double *bd = reinterpret_cast<double *>(buf);
    const int total = G * n;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int g, j;
        elem_of<AXIS>(e, n, G, g, j);
        LineGeom lg;
        double v = 0.0;
        if (line_of<AXIS>(grp, g, G, nx, ny, nz, lg)) {
            if constexpr (std::is_same<SrcT, unsigned long long>::value) {
                v = (double)(long long)src[lg.base + (int64_t)j * estride] * src_scale;
            } else {
                v = (double)src[lg.base + (int64_t)j * estride];
            }
        }
        bd[((g >> 1) * n + j) * 2 + (g & 1)] = v;
    }
}

// Halfcomplex value at position m of line (pair c, which) from the complex
// spectrum X of the packed pair.
__device__ __forceinline__ double hc_value(const cplx *X, int n, int c, int which, int m)
{
    const int k = (2 * m <= n) ? m : n - m;
    const cplx zk = X[c * n + k];
    const cplx znk = X[c * n + (k == 0 ? 0 : n - k)];
    if (which == 0) {
        return (2 * m <= n) ? 0.5 * (zk.re + znk.re) : 0.5 * (zk.im - znk.im);
    } else {
        return (2 * m <= n) ? 0.5 * (zk.im + znk.im) : 0.5 * (znk.re - zk.re);
    }
}

// Build the packed complex spectrum Z = A + i B from halfcomplex pairs H
// (H[c][m] = (hc_a[m], hc_b[m])).
__device__ void build_packed(const cplx *H, cplx *Z, int n, int nlines)
{
    const int total = n * nlines;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int c = t / n;
        const int k = t - c * n;
        const cplx hk = H[c * n + k];
        const cplx hnk = H[c * n + (k == 0 ? 0 : n - k)];
        double ra, ia, rb, ib;
        if (k == 0 || 2 * k == n) {
            ra = hk.re;
            ia = 0.0;
            rb = hk.im;
            ib = 0.0;
        } else if (2 * k < n) {
            ra = hk.re;
            ia = hnk.re;
            rb = hk.im;
            ib = hnk.im;
        } else {
            ra = hnk.re;
            ia = -hk.re;
            rb = hnk.im;
            ib = -hk.im;
        }
        Z[t] = {ra - ib, ia + rb};
    }
}

template <int KIND, int AXIS, typename SrcT>
__global__ void __launch_bounds__(kFftBlock)
    k_r2r_axis(const SrcT *__restrict__ src, double *__restrict__ dst, double src_scale,
               double out_scale, int nx, int ny, int nz, FftAxisDev ax, int G,
               const double *__restrict__ lamx, const double *__restrict__ lamy,
               const double *__restrict__ lamz)
{
    extern __shared__ cplx lds[];
    const int n = ax.n;
    const int nlines = (G + 1) / 2;
    cplx *A = lds;
    cplx *B = lds + nlines * n;
    const int grp = blockIdx.x;
    const int64_t estride = AXIS == 0 ? 1 : (AXIS == 1 ? (int64_t)nx : (int64_t)nx * ny);

    load_lines<AXIS, SrcT>(src, src_scale, A, grp, G, n, estride, nx, ny, nz);
    __syncthreads();

    cplx *R;
    if (KIND == kFftR2HC || KIND == kFftPoisson) {
        cplx *X = fft_lds<-1>(A, B, n, nlines, ax);
        if (KIND == kFftR2HC) {
            R = X;
        } else {
            // halfcomplex pairs, divided by the Green's-function denominator
            cplx *Y = (X == A) ? B : A;
            const int total = n * nlines;
            for (int t = threadIdx.x; t < total; t += blockDim.x) {
                const int c = t / n;
                const int m = t - c * n;
                double va = hc_value(X, n, c, 0, m);
                double vb = hc_value(X, n, c, 1, m);
                LineGeom la, lb;
                const bool oka = line_of<AXIS>(grp, 2 * c, G, nx, ny, nz, la);
                const bool okb = line_of<AXIS>(grp, 2 * c + 1, G, nx, ny, nz, lb);
                if (oka) {
                    const double lam = lamx[la.ix] + lamy[la.iy] + lamz[m];
                    va = (la.ix == 0 && la.iy == 0 && m == 0) ? 0.0 : va / lam;
                }
                if (okb) {
                    const double lam = lamx[lb.ix] + lamy[lb.iy] + lamz[m];
                    vb = (lb.ix == 0 && lb.iy == 0 && m == 0) ? 0.0 : vb / lam;
                }
                Y[t] = {va, vb};
            }
            __syncthreads();
            build_packed(Y, X, n, nlines);
            __syncthreads();
            R = fft_lds<+1>(X, Y, n, nlines, ax);
        }
    } else {
        build_packed(A, B, n, nlines);
        __syncthreads();
        R = fft_lds<+1>(B, A, n, nlines, ax);
    }

    // write back
    const int total = G * n;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int g, j;
        elem_of<AXIS>(e, n, G, g, j);
        LineGeom lg;
        if (!line_of<AXIS>(grp, g, G, nx, ny, nz, lg)) continue;
        double v;
        if (KIND == kFftR2HC) {
            v = hc_value(R, n, g >> 1, g & 1, j);
        } else {
            const cplx z = R[(g >> 1) * n + j];
            v = (g & 1) ? z.im : z.re;
        }
        dst[lg.base + (int64_t)j * estride] = v * out_scale;
    }
}

}  // namespace

int fft_group_lines(int n)
{
    // real lines per workgroup: keep 2 LDS buffers of G/2 complex lines of n
    // under 64 KiB, at least 2 lines, at most 64
    int g = 4096 / (n > 0 ? n : 1);
    if (g > 64) g = 64;
    if (g < 2) g = 2;
    return g & ~1;
}

size_t fft_lds_bytes(int n, int G)
{
    return (size_t)2 * ((G + 1) / 2) * n * sizeof(cplx);
}

int64_t fft_groups(int axis, int G, int nx, int ny, int nz)
{
    if (axis == 0) return ((int64_t)ny * nz + G - 1) / G;
    if (axis == 1) return (int64_t)((nx + G - 1) / G) * nz;
    return ((int64_t)nx * ny + G - 1) / G;
}

template <int KIND, int AXIS, typename SrcT>
static hipError_t launch_one(const SrcT *src, double *dst, double src_scale, double out_scale,
                             int nx, int ny, int nz, const FftAxisDev &ax, const double *lx,
                             const double *ly, const double *lz, hipStream_t st)
{
    const int G = fft_group_lines(ax.n);
    const int64_t groups = fft_groups(AXIS, G, nx, ny, nz);
    if (groups <= 0) return hipSuccess;
    const size_t lds = fft_lds_bytes(ax.n, G);
    hipLaunchKernelGGL((k_r2r_axis<KIND, AXIS, SrcT>), dim3((unsigned)groups), dim3(kFftBlock),
                       lds, st, src, dst, src_scale, out_scale, nx, ny, nz, ax, G, lx, ly, lz);
    return hipGetLastError();
}

hipError_t fft_axis_pass(int kind, int axis, const void *src, bool src_is_fixed,
                         double src_scale, double *dst, double out_scale, int nx, int ny,
                         int nz, const FftAxisDev &ax, const double *lamx, const double *lamy,
                         const double *lamz, hipStream_t st)
{
#define OHPIC_FFT_DISPATCH(K, A)                                                               \
    if (src_is_fixed)                                                                          \
        return launch_one<K, A, unsigned long long>(                                           \
            (const unsigned long long *)src, dst, src_scale, out_scale, nx, ny, nz, ax, lamx,  \
            lamy, lamz, st);                                                                   \
    return launch_one<K, A, double>((const double *)src, dst, src_scale, out_scale, nx, ny,    \
                                    nz, ax, lamx, lamy, lamz, st);
    if (kind == kFftR2HC) {
        if (axis == 0) { OHPIC_FFT_DISPATCH(kFftR2HC, 0) }
        if (axis == 1) { OHPIC_FFT_DISPATCH(kFftR2HC, 1) }
        OHPIC_FFT_DISPATCH(kFftR2HC, 2)
    } else if (kind == kFftHC2R) {
        if (axis == 0) { OHPIC_FFT_DISPATCH(kFftHC2R, 0) }
        if (axis == 1) { OHPIC_FFT_DISPATCH(kFftHC2R, 1) }
        OHPIC_FFT_DISPATCH(kFftHC2R, 2)
    } else {
        if (axis != 2) return hipErrorInvalidValue;
        OHPIC_FFT_DISPATCH(kFftPoisson, 2)
    }
#undef OHPIC_FFT_DISPATCH
}

}  // namespace ohpic
