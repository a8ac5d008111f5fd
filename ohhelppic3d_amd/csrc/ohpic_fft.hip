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


struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// One Stockham radix-P stage, in place over `nlines` complex lines of length
// n held in LDS.  Task (line, row, f): inputs y[(row + q*r)*Ls + f] times
// w_L^{f q}, outputs y[row*L + f + Ls*u] = sum_q (...) w_p^{u q}.  Every
// thread reads the inputs of all its tasks into registers, the block
// synchronises, then the outputs are written back over the same buffer.
template <int SIGN>
__device__ __forceinline__ cplx twiddle(const double2 *__restrict__ tw, int k)
{
    const double2 w = tw[k];
    return cplx{w.x, SIGN < 0 ? w.y : -w.y};
}

template <int SIGN, int P>
__device__ __forceinline__ void butterfly(cplx *a, const double2 *__restrict__ tw, int step)
{
    if (P == 2) {
        const cplx t = a[1];
        a[1] = {a[0].re - t.re, a[0].im - t.im};
        a[0] = {a[0].re + t.re, a[0].im + t.im};
    } else if (P == 4) {
        const cplx t0 = {a[0].re + a[2].re, a[0].im + a[2].im};
        const cplx t1 = {a[0].re - a[2].re, a[0].im - a[2].im};
        const cplx t2 = {a[1].re + a[3].re, a[1].im + a[3].im};
        const cplx t3 = {a[1].re - a[3].re, a[1].im - a[3].im};
        const cplx wt3 = {-SIGN * t3.im, SIGN * t3.re};  // (SIGN i) * t3
        a[0] = {t0.re + t2.re, t0.im + t2.im};
        a[1] = {t1.re + wt3.re, t1.im + wt3.im};
        a[2] = {t0.re - t2.re, t0.im - t2.im};
        a[3] = {t1.re - wt3.re, t1.im - wt3.im};
    } else if (P == 3) {
        const double sn = SIGN * 0.86602540378443864676372317075294;  // sin(2 pi/3)
        const cplx ts = {a[1].re + a[2].re, a[1].im + a[2].im};
        const cplx td = {a[1].re - a[2].re, a[1].im - a[2].im};
        const cplx m = {a[0].re - 0.5 * ts.re, a[0].im - 0.5 * ts.im};
        const cplx j = {-sn * td.im, sn * td.re};
        a[0] = {a[0].re + ts.re, a[0].im + ts.im};
        a[1] = {m.re + j.re, m.im + j.im};
        a[2] = {m.re - j.re, m.im - j.im};
    } else {
        cplx o[P];
#pragma unroll
        for (int u = 0; u < P; ++u) {
            cplx acc = a[0];
#pragma unroll
            for (int q = 1; q < P; ++q) {
                const cplx v = cmul(a[q], twiddle<SIGN>(tw, ((u * q) % P) * step));
                acc.re += v.re;
                acc.im += v.im;
            }
            o[u] = acc;
        }
#pragma unroll
        for (int u = 0; u < P; ++u) a[u] = o[u];
    }
}

template <int SIGN, int P>
__device__ void fft_stage(cplx *__restrict__ buf, int n, int nlines, int Ls,
                          const double2 *__restrict__ tw)
{
    constexpr int KMAX = (kFftLdsCplx / kFftBlock + P - 1) / P;  // tasks per thread
    const int L = Ls * P;
    const int r = n / L;
    const int tpl = n / P;
    const int total = tpl * nlines;
    cplx a[KMAX][P];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int t = threadIdx.x + k * kFftBlock;
        if (t < total) {
            const int line = t / tpl;
            const int i = t - line * tpl;
            const int row = i / Ls;
            const int f = i - row * Ls;
            const cplx *s = buf + line * n;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                cplx v = s[(row + q * r) * Ls + f];
                if (q > 0 && f > 0) v = cmul(v, twiddle<SIGN>(tw, f * q * r));
                a[k][q] = v;
            }
            butterfly<SIGN, P>(a[k], tw, n / P);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int t = threadIdx.x + k * kFftBlock;
        if (t < total) {
            const int line = t / tpl;
            const int i = t - line * tpl;
            const int row = i / Ls;
            const int f = i - row * Ls;
            cplx *d = buf + line * n + row * L + f;
#pragma unroll
            for (int u = 0; u < P; ++u) d[Ls * u] = a[k][u];
        }
    }
    __syncthreads();
}

// Runs all stages in place.  RS selects the compiled radix set (0: 4 and 2;
// 1: 4, 2 and 3; 2: every prime up to kMaxRadix) so power-of-two lengths do
// not pay the register cost of the large generic stages.
template <int SIGN, int RS>
__device__ void fft_lds(cplx *buf, int n, int nlines, const FftAxisDev &ax)
{
    int Ls = 1;
    for (int s = 0; s < ax.nfac; ++s) {
        const int p = ax.fac[s];
        if (p == 4) {
            fft_stage<SIGN, 4>(buf, n, nlines, Ls, ax.tw);
        } else if (p == 2) {
            fft_stage<SIGN, 2>(buf, n, nlines, Ls, ax.tw);
        } else if (RS >= 1 && p == 3) {
            fft_stage<SIGN, 3>(buf, n, nlines, Ls, ax.tw);
        } else if (RS >= 2) {
            switch (p) {
            case 5: fft_stage<SIGN, 5>(buf, n, nlines, Ls, ax.tw); break;
            case 7: fft_stage<SIGN, 7>(buf, n, nlines, Ls, ax.tw); break;
            case 11: fft_stage<SIGN, 11>(buf, n, nlines, Ls, ax.tw); break;
            default: fft_stage<SIGN, 13>(buf, n, nlines, Ls, ax.tw); break;
            }
        }
        Ls *= p;
    }
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

// Packed complex spectrum Z[k] = A[k] + i B[k] from halfcomplex pairs
// H[c][m] = (hc_a[m], hc_b[m]).
__device__ __forceinline__ cplx packed_from_hc(const cplx *H, int n, int c, int k)
{
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
    return cplx{ra - ib, ia + rb};
}

constexpr int kElemMax = kFftLdsCplx / kFftBlock;  // LDS elements per thread

// In place: buf <- packed spectrum built from the halfcomplex pairs in buf.
__device__ void build_packed_inplace(cplx *buf, int n, int nlines)
{
    const int total = n * nlines;
    cplx z[kElemMax];
#pragma unroll
    for (int k = 0; k < kElemMax; ++k) {
        const int t = threadIdx.x + k * kFftBlock;
        if (t < total) {
            const int c = t / n;
            z[k] = packed_from_hc(buf, n, c, t - c * n);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kElemMax; ++k) {
        const int t = threadIdx.x + k * kFftBlock;
        if (t < total) buf[t] = z[k];
    }
    __syncthreads();
}

template <int KIND, int AXIS, int RS, typename SrcT>
__global__ void __launch_bounds__(kFftBlock)
    k_r2r_axis(const SrcT *__restrict__ src, double *__restrict__ dst, double src_scale,
               double out_scale, int nx, int ny, int nz, FftAxisDev ax, int G,
               const double *__restrict__ lamx, const double *__restrict__ lamy,
               const double *__restrict__ lamz)
{
    __shared__ cplx buf[kFftLdsCplx];
    const int n = ax.n;
    const int nlines = (G + 1) / 2;
    const int grp = blockIdx.x;
    const int64_t estride = AXIS == 0 ? 1 : (AXIS == 1 ? (int64_t)nx : (int64_t)nx * ny);

    load_lines<AXIS, SrcT>(src, src_scale, buf, grp, G, n, estride, nx, ny, nz);
    __syncthreads();

    if (KIND == kFftR2HC || KIND == kFftPoisson) {
        fft_lds<-1, RS>(buf, n, nlines, ax);
        if (KIND == kFftPoisson) {
            // halfcomplex pairs divided by the Green's-function denominator
            const int total = n * nlines;
            cplx h[kElemMax];
#pragma unroll
            for (int k = 0; k < kElemMax; ++k) {
                const int t = threadIdx.x + k * kFftBlock;
                if (t < total) {
                    const int c = t / n;
                    const int m = t - c * n;
                    double va = hc_value(buf, n, c, 0, m);
                    double vb = hc_value(buf, n, c, 1, m);
                    LineGeom la, lb;
                    if (line_of<AXIS>(grp, 2 * c, G, nx, ny, nz, la)) {
                        const double lam = lamx[la.ix] + lamy[la.iy] + lamz[m];
                        va = (la.ix == 0 && la.iy == 0 && m == 0) ? 0.0 : va / lam;
                    }
                    if (line_of<AXIS>(grp, 2 * c + 1, G, nx, ny, nz, lb)) {
                        const double lam = lamx[lb.ix] + lamy[lb.iy] + lamz[m];
                        vb = (lb.ix == 0 && lb.iy == 0 && m == 0) ? 0.0 : vb / lam;
                    }
                    h[k] = {va, vb};
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kElemMax; ++k) {
                const int t = threadIdx.x + k * kFftBlock;
                if (t < total) buf[t] = h[k];
            }
            __syncthreads();
            build_packed_inplace(buf, n, nlines);
            fft_lds<+1, RS>(buf, n, nlines, ax);
        }
    } else {
        build_packed_inplace(buf, n, nlines);
        fft_lds<+1, RS>(buf, n, nlines, ax);
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
            v = hc_value(buf, n, g >> 1, g & 1, j);
        } else {
            const cplx z = buf[(g >> 1) * n + j];
            v = (g & 1) ? z.im : z.re;
        }
        dst[lg.base + (int64_t)j * estride] = v * out_scale;
    }
}

}  // namespace

int fft_group_lines(int n)
{
    // real lines per workgroup: G/2 complex lines of n fill at most
    // kFftLdsCplx complex values (32 KiB); at least 2 lines, at most 64
    int g = 2 * (kFftLdsCplx / (n > 0 ? n : 1));
    if (g > 64) g = 64;
    if (g < 2) g = 2;
    return g & ~1;
}

size_t fft_lds_bytes(int n, int G)
{
    return (size_t)((G + 1) / 2) * n * sizeof(cplx);
}

int64_t fft_groups(int axis, int G, int nx, int ny, int nz)
{
    if (axis == 0) return ((int64_t)ny * nz + G - 1) / G;
    if (axis == 1) return (int64_t)((nx + G - 1) / G) * nz;
    return ((int64_t)nx * ny + G - 1) / G;
}

int fft_radix_set(const FftAxisDev &ax)
{
    int rs = 0;
    for (int i = 0; i < ax.nfac; ++i) {
        if (ax.fac[i] == 3 && rs < 1) rs = 1;
        if (ax.fac[i] > 4) rs = 2;
    }
    return rs;
}

template <int KIND, int AXIS, typename SrcT>
static hipError_t launch_one(const SrcT *src, double *dst, double src_scale, double out_scale,
                             int nx, int ny, int nz, const FftAxisDev &ax, const double *lx,
                             const double *ly, const double *lz, hipStream_t st)
{
    const int G = fft_group_lines(ax.n);
    const int64_t groups = fft_groups(AXIS, G, nx, ny, nz);
    if (groups <= 0) return hipSuccess;
    const dim3 g((unsigned)groups), b(kFftBlock);
    switch (fft_radix_set(ax)) {
    case 0:
        hipLaunchKernelGGL((k_r2r_axis<KIND, AXIS, 0, SrcT>), g, b, 0, st, src, dst, src_scale,
                           out_scale, nx, ny, nz, ax, G, lx, ly, lz);
        break;
    case 1:
        hipLaunchKernelGGL((k_r2r_axis<KIND, AXIS, 1, SrcT>), g, b, 0, st, src, dst, src_scale,
                           out_scale, nx, ny, nz, ax, G, lx, ly, lz);
        break;
    default:
        hipLaunchKernelGGL((k_r2r_axis<KIND, AXIS, 2, SrcT>), g, b, 0, st, src, dst, src_scale,
                           out_scale, nx, ny, nz, ax, G, lx, ly, lz);
        break;
    }
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
