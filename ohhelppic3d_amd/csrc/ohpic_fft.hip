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
#include "ohpic_internal.hpp"

namespace ohpic {

namespace {


struct cplx {
    double re, im;
};

__device__ __forceinline__ int fastdiv(int t, uint32_t m)
{
    return m ? (int)__umulhi((uint32_t)t, m) : t;
}

__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// One Stockham radix-P stage, in place over `nlines` complex lines of length
// n held in LDS.  Task (line, row, f): inputs y[(row + q*r)*Ls + f] times
// w_L^{f q}, outputs y[row*L + f + Ls*u] = sum_q (...) w_p^{u q}.  Every
// thread reads the inputs of all its tasks into registers, the block
// synchronises, then the outputs are written back over the same buffer.
// Twiddle tables live in global memory, or (TL) in the workgroup's LDS copy
// for lengths up to kTwLds: staged once per workgroup (4 KiB at n = 256, so
// 8 workgroups still fit a CU), they cut every single-transform pass at 256^3
// from 60-62 us to 56-59 us.
constexpr int kTwLds = 256;
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) d2v lds_d2v;

template <int SIGN, bool TL>
__device__ __forceinline__ cplx twiddle(const double2 *__restrict__ tw, int k)
{
    // one 16-byte load (a native vector type: HIP's double2 class does not
    // bind to an address-space-qualified reference)
    d2v w;
    if (TL)
        w = ((const lds_d2v *)tw)[k];
    else
        w = gload(reinterpret_cast<const d2v *>(tw), k);
    return cplx{w.x, SIGN < 0 ? w.y : -w.y};
}

template <int SIGN, int P, bool TL>
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
                const cplx v = cmul(a[q], twiddle<SIGN, TL>(tw, ((u * q) % P) * step));
                acc.re += v.re;
                acc.im += v.im;
            }
            o[u] = acc;
        }
#pragma unroll
        for (int u = 0; u < P; ++u) a[u] = o[u];
    }
}

template <int SIGN, int P, bool TL>
__device__ __forceinline__ void fft_stage(cplx *__restrict__ buf, int n, int nlines, int Ls,
                                          const double2 *__restrict__ tw, uint32_t mg_tpl,
                                          uint32_t mg_Ls)
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
            const int line = fastdiv(t, mg_tpl);
            const int i = t - line * tpl;
            const int row = fastdiv(i, mg_Ls);
            const int f = i - row * Ls;
            const cplx *s = buf + line * n;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                cplx v = s[(row + q * r) * Ls + f];
                if (q > 0 && f > 0) v = cmul(v, twiddle<SIGN, TL>(tw, f * q * r));
                a[k][q] = v;
            }
            butterfly<SIGN, P, TL>(a[k], tw, n / P);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int t = threadIdx.x + k * kFftBlock;
        if (t < total) {
            const int line = fastdiv(t, mg_tpl);
            const int i = t - line * tpl;
            const int row = fastdiv(i, mg_Ls);
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
template <int SIGN, int RS, bool TL>
__device__ __forceinline__ void fft_lds_t(cplx *buf, int n, int nlines, const FftAxisDev &ax,
                                          const double2 *tw)
{
    int Ls = 1;
    for (int s = 0; s < ax.nfac; ++s) {
        const int p = ax.fac[s];
        const uint32_t mt = ax.mg_tpl[s], ml = ax.mg_Ls[s];
        if (p == 4) {
            fft_stage<SIGN, 4, TL>(buf, n, nlines, Ls, tw, mt, ml);
        } else if (p == 2) {
            fft_stage<SIGN, 2, TL>(buf, n, nlines, Ls, tw, mt, ml);
        } else if (RS >= 1 && p == 3) {
            fft_stage<SIGN, 3, TL>(buf, n, nlines, Ls, tw, mt, ml);
        } else if (RS >= 2) {
            switch (p) {
            case 5: fft_stage<SIGN, 5, TL>(buf, n, nlines, Ls, tw, mt, ml); break;
            case 7: fft_stage<SIGN, 7, TL>(buf, n, nlines, Ls, tw, mt, ml); break;
            case 11: fft_stage<SIGN, 11, TL>(buf, n, nlines, Ls, tw, mt, ml); break;
            default: fft_stage<SIGN, 13, TL>(buf, n, nlines, Ls, tw, mt, ml); break;
            }
        }
        Ls *= p;
    }
}

// All stages, twiddles from the LDS copy stw when the kernel staged them
// (lengths up to kTwLds), else from the global table.
template <int SIGN, int RS>
__device__ __forceinline__ void fft_lds(cplx *buf, int n, int nlines, const FftAxisDev &ax,
                                        const double2 *stw)
{
    if (n <= kTwLds)
        fft_lds_t<SIGN, RS, true>(buf, n, nlines, ax, stw);
    else
        fft_lds_t<SIGN, RS, false>(buf, n, nlines, ax, ax.tw);
}

// The G lines of one workgroup: line g (< nvalid) starts at element
// base + g * gstride and its elements are estride apart.  x: G consecutive
// rows; y: G consecutive x positions of one z plane; z: G consecutive (x, y)
// positions.
struct LineBlock {
    int64_t base;
    int64_t estride;
    int gstride;
    int nvalid;
};

template <int AXIS>
__device__ __forceinline__ LineBlock line_block(int grp, int G, int nx, int ny, int nz)
{
    LineBlock b;
    if (AXIS == 0) {
        const int64_t row0 = (int64_t)grp * G;
        b.base = row0 * nx;
        b.gstride = nx;
        b.estride = 1;
        b.nvalid = (int)min<int64_t>(G, (int64_t)ny * nz - row0);
    } else if (AXIS == 1) {
        const int gpp = (nx + G - 1) / G;
        const int iz = grp / gpp;
        const int ix0 = (grp - iz * gpp) * G;
        b.base = ix0 + (int64_t)nx * ny * iz;
        b.gstride = 1;
        b.estride = nx;
        b.nvalid = min(G, nx - ix0);
    } else {
        const int64_t l0 = (int64_t)grp * G;
        b.base = l0;
        b.gstride = 1;
        b.estride = (int64_t)nx * ny;
        b.nvalid = (int)min<int64_t>(G, (int64_t)nx * ny - l0);
    }
    return b;
}

// element e of the group -> (line g, position j): x lines are read row by
// row (g = e / n), y and z lines across the G lines first (j = e / G) so a
// wavefront touches G consecutive doubles per position
template <int AXIS>
__device__ __forceinline__ void elem_of(int e, int n, int G, const FftAxisDev &ax, int &g, int &j)
{
    if (AXIS == 0) {
        g = fastdiv(e, ax.mg_n);
        j = e - g * n;
    } else {
        j = fastdiv(e, ax.mg_G);
        g = e - j * G;
    }
}

constexpr int kLineElems = 2 * kFftLdsCplx / kFftBlock;  // real elements per thread

// Load G real lines into LDS packed as complex (line pair -> re/im).  Every
// thread issues all of its (up to kLineElems) global loads before the first
// LDS write, so each lane keeps that many reads in flight.
template <int AXIS>
__device__ void load_lines(const gmem_t<double> *__restrict__ src, cplx *buf, const LineBlock &lb,
                           int G, int n, const FftAxisDev &ax)
{
    double *bd = reinterpret_cast<double *>(buf);
    const int total = G * n;
    double v[kLineElems];
    int at[kLineElems];
#pragma unroll
    for (int k = 0; k < kLineElems; ++k) {
        const int e = threadIdx.x + k * kFftBlock;
        v[k] = 0.0;
        at[k] = -1;
        if (e < total) {
            int g, j;
            elem_of<AXIS>(e, n, G, ax, g, j);
            if (g < lb.nvalid)
                v[k] = src[lb.base + (int64_t)g * lb.gstride + (int64_t)j * lb.estride];
            at[k] = ((g >> 1) * n + j) * 2 + (g & 1);
        }
    }
#pragma unroll
    for (int k = 0; k < kLineElems; ++k)
        if (at[k] >= 0) bd[at[k]] = v[k];
}


// Task t of a (line pair c, position k < per_c) sweep: x lines take
// consecutive positions of one pair per wavefront (rows stay contiguous), y
// and z lines consecutive pairs at one position (the G lines of a position are
// adjacent in memory).
template <int AXIS>
__device__ __forceinline__ void pair_task(int t, int per_c, uint32_t mg_per, int nlines,
                                          uint32_t mg_nl, int &c, int &k)
{
    if (AXIS == 0) {
        c = fastdiv(t, mg_per);
        k = t - c * per_c;
    } else {
        k = fastdiv(t, mg_nl);
        c = t - k * nlines;
    }
}

// R2HC output: the task (c, k <= n/2) reads Z_k and Z_{n-k} once and writes
// r_k (position k) and i_k (position n - k) of both lines of the pair, with
// A_k = (Z_k + conj Z_{n-k}) / 2 and B_k = (Z_k - conj Z_{n-k}) / 2i.
template <int AXIS>
__device__ __forceinline__ void store_r2hc_pairs(const cplx *buf, gmem_t<double> *__restrict__ dst,
                                                 const LineBlock &lb, int n, int nlines,
                                                 const FftAxisDev &ax)
{
    const int half = n / 2 + 1;
    const int total = half * nlines;
    for (int t = threadIdx.x; t < total; t += kFftBlock) {
        int c, k;
        pair_task<AXIS>(t, half, ax.mg_half, nlines, ax.mg_nl, c, k);
        const int km = k == 0 ? 0 : n - k;
        const cplx zk = buf[c * n + k], zm = buf[c * n + km];
        const double ar = 0.5 * (zk.re + zm.re), ai = 0.5 * (zk.im - zm.im);
        const double br = 0.5 * (zk.im + zm.im), bi = 0.5 * (zm.re - zk.re);
        const int64_t la = lb.base + (int64_t)(2 * c) * lb.gstride;
        if (2 * c < lb.nvalid) {
            dst[la + (int64_t)k * lb.estride] = ar;
            if (km != k) dst[la + (int64_t)km * lb.estride] = ai;
        }
        if (2 * c + 1 < lb.nvalid) {
            const int64_t lbb = la + lb.gstride;
            dst[lbb + (int64_t)k * lb.estride] = br;
            if (km != k) dst[lbb + (int64_t)km * lb.estride] = bi;
        }
    }
}

// HC2R input: the task (c, k <= n/2) loads positions k and n - k of both
// halfcomplex lines and writes the packed spectrum Z = A + i B at k and n - k
// (A_k = r_k + i i_k from line a, B_k likewise from line b).
template <int AXIS>
__device__ __forceinline__ void load_hc_packed(const gmem_t<double> *__restrict__ src, cplx *buf,
                                               const LineBlock &lb, int n, int nlines,
                                               const FftAxisDev &ax)
{
    const int half = n / 2 + 1;
    const int total = half * nlines;
    for (int t = threadIdx.x; t < total; t += kFftBlock) {
        int c, k;
        pair_task<AXIS>(t, half, ax.mg_half, nlines, ax.mg_nl, c, k);
        const int km = k == 0 ? 0 : n - k;
        const int64_t la = lb.base + (int64_t)(2 * c) * lb.gstride;
        double ak = 0.0, am = 0.0, bk = 0.0, bm = 0.0;
        if (2 * c < lb.nvalid) {
            ak = src[la + (int64_t)k * lb.estride];
            am = src[la + (int64_t)km * lb.estride];
        }
        if (2 * c + 1 < lb.nvalid) {
            bk = src[la + lb.gstride + (int64_t)k * lb.estride];
            bm = src[la + lb.gstride + (int64_t)km * lb.estride];
        }
        if (km == k) {  // k = 0 and k = n/2: real r_k of both lines
            buf[c * n + k] = cplx{ak, bk};
        } else {
            buf[c * n + k] = cplx{ak - bm, am + bk};
            buf[c * n + km] = cplx{ak + bm, bk - am};
        }
    }
}

// Complex output (HC2R and the Poisson pass): position m of the pair's
// spectrum holds (line a, line b) = (re, im).
template <int AXIS>
__device__ __forceinline__ void store_complex_pairs(const cplx *buf, gmem_t<double> *__restrict__ dst,
                                                    const LineBlock &lb, int n, int nlines,
                                                    const FftAxisDev &ax, double scale)
{
    const int total = n * nlines;
    for (int t = threadIdx.x; t < total; t += kFftBlock) {
        int c, m;
        pair_task<AXIS>(t, n, ax.mg_n, nlines, ax.mg_nl, c, m);
        const cplx z = buf[c * n + m];
        const int64_t la = lb.base + (int64_t)(2 * c) * lb.gstride + (int64_t)m * lb.estride;
        if (2 * c < lb.nvalid) dst[la] = z.re * scale;
        if (2 * c + 1 < lb.nvalid) dst[la + lb.gstride] = z.im * scale;
    }
}

// One axis pass for every job of the table (blockIdx.y = job, blockIdx.x =
// line group; groups beyond a job's count exit).  FROM_SRC reads job.src
// (rho) instead of transforming job.data in place; SCALE multiplies the
// output by job.inv_n (the final HC2R pass).
#ifndef OHPIC_FFT_MIN_WAVES
#define OHPIC_FFT_MIN_WAVES 8
#endif

template <int KIND, int AXIS, int RS, bool FROM_SRC, bool SCALE>
__global__ void __launch_bounds__(kFftBlock)
    __attribute__((amdgpu_waves_per_eu(OHPIC_FFT_MIN_WAVES, 8)))
    k_fft_pass(const FftJob *__restrict__ jobs, int skip256)
{
    __shared__ cplx buf[kFftLdsCplx];
    __shared__ double sLamXY[KIND == kFftPoisson ? 64 : 1];
    __shared__ double2 sTw[kTwLds];
    const FftJob &J = jobs[blockIdx.y];
    const int nx = J.n[0], ny = J.n[1], nz = J.n[2];
    const FftAxisDev &ax = J.ax[AXIS];
    const int n = ax.n;
    // a Poisson job with 256-point z lines is k_poisson_z256's
    if (KIND == kFftPoisson && skip256 && n == 256) return;
    const int G = fft_group_lines(n);
    const int64_t ngroups = fft_groups(AXIS, G, nx, ny, nz);
    if ((int64_t)blockIdx.x >= ngroups) return;
    // neighbouring line groups read the two halves of the same 128-byte
    // lines (y and z passes): keep them on one XCD
    const int grp = xcd_remap((int)blockIdx.x, (int)ngroups);
    const gmem_t<double> *__restrict__ src = as_global(FROM_SRC ? J.src : J.data);
    gmem_t<double> *__restrict__ dst = as_global(J.data);
    const double out_scale = SCALE ? J.inv_n : 1.0;
    const int nlines = (G + 1) / 2;
    const LineBlock lb = line_block<AXIS>(grp, G, nx, ny, nz);

    if (KIND == kFftPoisson && (int)threadIdx.x < G) {
        // lam_x(ix) + lam_y(iy) of z line g; the DC line (ix = iy = 0) is marked
        // by -1 (its k_z = 0 mode is set to zero)
        const int g = threadIdx.x;
        double v = 0.0;
        if (g < lb.nvalid) {
            const int l = (int)lb.base + g;
            const int iy = l / nx;
            const int ix = l - iy * nx;
            v = (l == 0) ? -1.0 : gload(J.lam[0], ix) + gload(J.lam[1], iy);
        }
        sLamXY[g] = v;
    }
    if (n <= kTwLds)
        for (int i = threadIdx.x; i < n; i += kFftBlock) {
            const d2v w = gload(reinterpret_cast<const d2v *>(ax.tw), i);
            sTw[i] = make_double2(w.x, w.y);
        }
    if (KIND == kFftHC2R)
        load_hc_packed<AXIS>(src, buf, lb, n, nlines, ax);
    else
        load_lines<AXIS>(src, buf, lb, G, n, ax);
    __syncthreads();

    if (KIND == kFftR2HC || KIND == kFftPoisson) {
        fft_lds<-1, RS>(buf, n, nlines, ax, sTw);
        if (KIND == kFftPoisson) {
            // Green's-function division on the packed spectra: with
            // Z = FFT(a + i b), A_k = (Z_k + conj Z_{n-k}) / 2 and
            // B_k = (Z_k - conj Z_{n-k}) / 2i; A and B are divided by their
            // lines' denominators (DC := 0) and repacked as A' + i B', whose
            // inverse transform is (n a', n b') as HC2R of the halfcomplex
            // quotients gives.  One task per (pair, k <= n/2) owns slots k and
            // n - k, so the update needs no barrier of its own; one reciprocal
            // per denominator (within an ulp of the quotient).
            const gmem_t<double> *__restrict__ lamz = as_global(J.lam[2]);
            const int half = n / 2 + 1;
            const int total = half * nlines;
            constexpr int kTasks = (kFftLdsCplx / 2 + kFftLdsCplx / 16 + kFftBlock - 1) / kFftBlock + 1;
            const uint32_t mg_half = ax.mg_half;
#pragma unroll
            for (int q = 0; q < kTasks; ++q) {
                const int t = threadIdx.x + q * kFftBlock;
                if (t >= total) break;
                const int c = fastdiv(t, mg_half);
                const int k = t - c * half;
                const int km = k == 0 ? 0 : n - k;
                const cplx zk = buf[c * n + k], zm = buf[c * n + km];
                cplx A = {0.5 * (zk.re + zm.re), 0.5 * (zk.im - zm.im)};
                cplx B = {0.5 * (zk.im + zm.im), -0.5 * (zk.re - zm.re)};
                const double la = sLamXY[2 * c], lbv = sLamXY[2 * c + 1];
                const double lz = lamz[k];
                if (2 * c < lb.nvalid) {
                    if (la < 0.0 && k == 0) {
                        A = {0.0, 0.0};
                    } else {
                        const double d = (la < 0.0 ? 0.0 : la) + lz;
                        const double rd = 1.0 / d;
                        A = {A.re * rd, A.im * rd};
                    }
                }
                if (2 * c + 1 < lb.nvalid) {
                    if (lbv < 0.0 && k == 0) {
                        B = {0.0, 0.0};
                    } else {
                        const double d = (lbv < 0.0 ? 0.0 : lbv) + lz;
                        const double rd = 1.0 / d;
                        B = {B.re * rd, B.im * rd};
                    }
                }
                buf[c * n + k] = cplx{A.re - B.im, A.im + B.re};
                if (km != k) buf[c * n + km] = cplx{A.re + B.im, B.re - A.im};
            }
            __syncthreads();
            fft_lds<+1, RS>(buf, n, nlines, ax, sTw);
        }
    } else {
        fft_lds<+1, RS>(buf, n, nlines, ax, sTw);
    }

    if (KIND == kFftR2HC)
        store_r2hc_pairs<AXIS>(buf, dst, lb, n, nlines, ax);
    else
        store_complex_pairs<AXIS>(buf, dst, lb, n, nlines, ax, out_scale);
}

// ------------------------------------------------ Poisson z pass, n = 256
// The fused z pass for z lines of 256 points as a 16 x 16 decomposition with
// the 16-point DFTs in registers: n = 16 n1 + n2, k = k1 + 16 k2.  A 64-thread
// workgroup holds four complex lines (8 real z lines at consecutive x, packed
// in pairs), 16 threads per line, 16 points per thread.  Forward: thread n2
// loads x[16 n1 + n2] (n1 = 0..15), DFT16 over n1, twiddle W256^{n2 k1},
// transpose through LDS, DFT16 over n2: thread k1 holds X[k1 + 16 k2].  The
// Green's-function division pairs X[k] with X[256 - k] (read back from LDS),
// the inverse runs the same steps backwards, and thread n2 stores x[16 n1 +
// n2].  Three LDS exchanges per visit against ten for the radix-4 stages of
// k_fft_pass (67 us against 105 us at 256^3).  Rows of the LDS exchange
// buffer are padded to 17 entries so that 16 lanes of one line cover all
// banks.
constexpr int kZ256Lines = 4;                 // complex lines per workgroup
constexpr int kZ256Threads = 16 * kZ256Lines;
constexpr int kZ256Row = 17;                  // padded row of 16 entries
constexpr int kZ256LineStride = 16 * kZ256Row;

// cos and sin of 2 pi j / 16, j = 0..15
__device__ __forceinline__ cplx w16(int j, int sign)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173,
                     r2 = 0.70710678118654752440;
    const double C[16] = {1.0, c1, r2, s1, 0.0, -s1, -r2, -c1, -1.0, -c1, -r2, -s1, 0.0, s1, r2, c1};
    const double S[16] = {0.0, s1, r2, c1, 1.0, c1, r2, s1, 0.0, -s1, -r2, -c1, -1.0, -c1, -r2, -s1};
    return cplx{C[j & 15], sign < 0 ? -S[j & 15] : S[j & 15]};
}

// In-place 16-point DFT of a[0..15] with exponent sign SIGN (-1 forward):
// 4-point DFTs over a[m + 4 j], twiddles w16^{m q}, 4-point DFTs over m;
// output X[q + 4 r] in a[q + 4 r].
template <int SIGN>
__device__ __forceinline__ void dft16(cplx *a)
{
    cplx t[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        cplx v[4] = {a[m], a[m + 4], a[m + 8], a[m + 12]};
        butterfly<SIGN, 4, false>(v, nullptr, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) t[m][q] = (m == 0 || q == 0) ? v[q] : cmul(v[q], w16(m * q, SIGN));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cplx v[4] = {t[0][q], t[1][q], t[2][q], t[3][q]};
        butterfly<SIGN, 4, false>(v, nullptr, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) a[q + 4 * r] = v[r];
    }
}

// 1/d for 0 < d < 2^1000 (d = 0 gives inf, which the caller discards): the
// hardware estimate (about 23 correct bits) refined by two Newton steps, so
// the error is that of the final rounding.  Replacing the 32 correctly
// rounded divisions per thread (and their branches) took the pass from 82 to
// 67 us at 256^3.
__device__ __forceinline__ double z256_recip(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = __fma_rn(__fma_rn(-d, r, 1.0), r, r);
    return __fma_rn(__fma_rn(-d, r, 1.0), r, r);
}

__global__ void __launch_bounds__(kZ256Threads) k_poisson_z256(const FftJob *__restrict__ jobs)
{
    __shared__ cplx xb[kZ256Lines * kZ256LineStride];
    __shared__ double2 sTw[256];
    __shared__ double sLam[2 * kZ256Lines];
    const FftJob &J = jobs[blockIdx.y];
    if (J.n[2] != 256) return;
    const int nx = J.n[0], ny = J.n[1];
    const int64_t nxy = (int64_t)nx * ny;
    const int ngroups = (int)((nxy + 2 * kZ256Lines - 1) / (2 * kZ256Lines));
    if ((int)blockIdx.x >= ngroups) return;
    // neighbouring groups read the two 64-byte halves of the same 128-byte
    // lines: keep them on one XCD
    const int grp = xcd_remap((int)blockIdx.x, ngroups);
    const int64_t l0 = (int64_t)grp * (2 * kZ256Lines);
    const int nvalid = (int)min<int64_t>(2 * kZ256Lines, nxy - l0);
    const int tid = threadIdx.x;
    const int c = tid >> 4, t = tid & 15;
    gmem_t<double> *__restrict__ data = as_global(J.data);
    const FftAxisDev &ax = J.ax[2];
    for (int i = tid; i < 256; i += kZ256Threads) {
        const d2v w = gload(reinterpret_cast<const d2v *>(ax.tw), i);
        sTw[i] = make_double2(w.x, w.y);
    }
    if (tid < 2 * kZ256Lines) {
        // lam_x + lam_y of real line tid; -1 marks the DC line (ix = iy = 0);
        // lines past the grid's end get 1 (they hold zeros; no 0/0)
        double v = 1.0;
        if (tid < nvalid) {
            const int64_t l = l0 + tid;
            const int iy = (int)(l / nx), ix = (int)(l - (int64_t)iy * nx);
            v = (l == 0) ? -1.0 : gload(J.lam[0], ix) + gload(J.lam[1], iy);
        }
        sLam[tid] = v;
    }
    // real lines 2c (re) and 2c + 1 (im): one 16-byte load per point when
    // both are present and nx*ny is even (l0 is even, so every pair is then
    // 16-byte aligned), else two 8-byte loads
    const bool va = 2 * c < nvalid, vb = 2 * c + 1 < nvalid;
    const bool v16 = vb && (nxy & 1) == 0;
    const int64_t la = l0 + 2 * c;
    cplx a[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
        const int64_t o = la + (int64_t)(16 * n1 + t) * nxy;
        if (v16) {
            const d2v v = ((const gmem_t<d2v> *)data)[o >> 1];
            a[n1] = cplx{v.x, v.y};
        } else {
            a[n1] = cplx{va ? data[o] : 0.0, vb ? data[o + 1] : 0.0};
        }
    }
    __syncthreads();  // sTw, sLam
    const lds_d2v *tw = (const lds_d2v *)sTw;
    cplx *xl = xb + c * kZ256LineStride;

    // forward: DFT16 over n1 (thread n2 = t), twiddle, transpose
    dft16<-1>(a);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
        cplx v = a[k1];
        if (k1 > 0 && t > 0) {
            const d2v w = tw[t * k1];
            v = cmul(v, cplx{w.x, w.y});
        }
        xl[t * kZ256Row + k1] = v;
    }
    __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) a[n2] = xl[n2 * kZ256Row + t];
    // DFT16 over n2 (thread k1 = t): a[k2] = X[t + 16 k2]
    dft16<-1>(a);
    __syncthreads();  // transpose reads done before the buffer is reused
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) xl[k2 * kZ256Row + t] = a[k2];
    __syncthreads();

    // Green's-function division of the packed spectra (as k_fft_pass):
    // A_k = (Z_k + conj Z_{n-k}) / 2, B_k = (Z_k - conj Z_{n-k}) / 2i, divided
    // by their lines' denominators (DC := 0), repacked as A' + i B'
    const double lam_a = sLam[2 * c], lam_b = sLam[2 * c + 1];
    const double lxa = lam_a < 0.0 ? 0.0 : lam_a, lxb = lam_b < 0.0 ? 0.0 : lam_b;
    const gmem_t<double> *__restrict__ lamz = as_global(J.lam[2]);
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
        const int k = t + 16 * k2;
        const int m1 = (16 - t) & 15;
        const int m2 = t == 0 ? ((16 - k2) & 15) : 15 - k2;
        const cplx zk = a[k2], zm = xl[m2 * kZ256Row + m1];
        const double lz = lamz[k];
        const double ra = z256_recip(lxa + lz), rb = z256_recip(lxb + lz);
        // the DC line's k = 0 mode is set to zero (its denominator is 0)
        const double sa = (lam_a < 0.0 && k == 0) ? 0.0 : 0.5 * ra;
        const double sb = (lam_b < 0.0 && k == 0) ? 0.0 : 0.5 * rb;
        const cplx A = {sa * (zk.re + zm.re), sa * (zk.im - zm.im)};
        const cplx B = {sb * (zk.im + zm.im), -sb * (zk.re - zm.re)};
        a[k2] = cplx{A.re - B.im, A.im + B.re};
    }

    // inverse: DFT16 over k2 (thread k1 = t), twiddle, transpose, DFT16 over k1
    dft16<+1>(a);
    __syncthreads();  // pair reads done before the buffer is reused
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
        cplx v = a[n2];
        if (n2 > 0 && t > 0) {
            const d2v w = tw[t * n2];
            v = cmul(v, cplx{w.x, -w.y});
        }
        xl[t * kZ256Row + n2] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) a[k1] = xl[k1 * kZ256Row + t];
    dft16<+1>(a);
    // thread n2 = t holds x[16 n1 + n2] in a[n1]
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
        const int64_t o = la + (int64_t)(16 * n1 + t) * nxy;
        if (v16) {
            d2v v;
            v.x = a[n1].re;
            v.y = a[n1].im;
            ((gmem_t<d2v> *)data)[o >> 1] = v;
        } else {
            if (va) data[o] = a[n1].re;
            if (vb) data[o + 1] = a[n1].im;
        }
    }
}

}  // namespace

void fft_plan_divisors(FftAxisDev &ax)
{
    ax.mg_n = fastdiv_magic((uint32_t)ax.n);
    ax.mg_G = fastdiv_magic((uint32_t)fft_group_lines(ax.n));
    ax.mg_half = fastdiv_magic((uint32_t)(ax.n / 2 + 1));
    ax.mg_nl = fastdiv_magic((uint32_t)(fft_group_lines(ax.n) / 2));
    int Ls = 1;
    for (int s = 0; s < kMaxFactors; ++s) {
        const int p = s < ax.nfac ? ax.fac[s] : 1;
        ax.mg_tpl[s] = fastdiv_magic((uint32_t)(ax.n / p));
        ax.mg_Ls[s] = fastdiv_magic((uint32_t)Ls);
        Ls *= p;
    }
}

int fft_radix_set(const FftAxisDev &ax)
{
    int rs = 0;
    for (int i = 0; i < ax.nfac; ++i) {
        if (ax.fac[i] == 3 && rs < 1) rs = 1;
        if (ax.fac[i] > 4 && ax.fac[i] != 3) rs = 2;
    }
    return rs;
}

template <int KIND, int AXIS, bool FROM_SRC, bool SCALE>
static hipError_t launch_pass(const FftJob *jobs, int njobs, int64_t max_groups, int rs,
                              hipStream_t st, int skip256 = 0)
{
    if (njobs <= 0 || max_groups <= 0) return hipSuccess;
    const dim3 g((unsigned)round_up_xcds(max_groups), (unsigned)njobs), b(kFftBlock);
    if (rs == 0)
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 0, FROM_SRC, SCALE>), g, b, 0, st, jobs, skip256);
    else if (rs == 1)
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 1, FROM_SRC, SCALE>), g, b, 0, st, jobs, skip256);
    else
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 2, FROM_SRC, SCALE>), g, b, 0, st, jobs, skip256);
    return hipGetLastError();
}

hipError_t fft_poisson_z(const FftJob *jobs, int njobs, int64_t max_groups, int rs, int z256,
                         hipStream_t st)
{
    if (njobs <= 0 || max_groups <= 0) return hipSuccess;
    if (z256 != kZ256All) {
        const hipError_t e = launch_pass<kFftPoisson, 2, false, false>(
            jobs, njobs, max_groups, rs, st, z256 == kZ256Some ? 1 : 0);
        if (e != hipSuccess || z256 == kZ256None) return e;
    }
    // the groups of a 256-point job are those of k_fft_pass (8 real lines each),
    // so max_groups covers them
    const dim3 g((unsigned)round_up_xcds(max_groups), (unsigned)njobs), b(kZ256Threads);
    hipLaunchKernelGGL(k_poisson_z256, g, b, 0, st, jobs);
    return hipGetLastError();
}

hipError_t fft_pass(FftPass pass, const FftJob *jobs, int njobs, int64_t max_groups, int rs,
                    hipStream_t st)
{
    switch (pass) {
    case kPassR2hcXFromSrc: return launch_pass<kFftR2HC, 0, true, false>(jobs, njobs, max_groups, rs, st);
    case kPassR2hcX: return launch_pass<kFftR2HC, 0, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassR2hcY: return launch_pass<kFftR2HC, 1, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassR2hcZ: return launch_pass<kFftR2HC, 2, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassPoissonZ: return launch_pass<kFftPoisson, 2, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassHc2rZ: return launch_pass<kFftHC2R, 2, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassHc2rY: return launch_pass<kFftHC2R, 1, false, false>(jobs, njobs, max_groups, rs, st);
    case kPassHc2rXScaled: return launch_pass<kFftHC2R, 0, false, true>(jobs, njobs, max_groups, rs, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace ohpic
