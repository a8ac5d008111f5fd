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
template <int AXIS>
__device__ void load_lines(const double *__restrict__ src, cplx *buf, int grp, int G, int n,
                           int64_t estride, int nx, int ny, int nz)
{
    double *bd = reinterpret_cast<double *>(buf);
    const int total = G * n;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int g, j;
        elem_of<AXIS>(e, n, G, g, j);
        LineGeom lg;
        double v = 0.0;
        if (line_of<AXIS>(grp, g, G, nx, ny, nz, lg)) v = src[lg.base + (int64_t)j * estride];
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

// One axis pass for every job of the table (blockIdx.y = job, blockIdx.x =
// line group; groups beyond a job's count exit).  FROM_SRC reads job.src
// (rho) instead of transforming job.data in place; SCALE multiplies the
// output by job.inv_n (the final HC2R pass).
template <int KIND, int AXIS, int RS, bool FROM_SRC, bool SCALE>
__global__ void __launch_bounds__(kFftBlock) k_fft_pass(const FftJob *__restrict__ jobs)
{
    __shared__ cplx buf[kFftLdsCplx];
    const FftJob &J = jobs[blockIdx.y];
    const int nx = J.n[0], ny = J.n[1], nz = J.n[2];
    const FftAxisDev &ax = J.ax[AXIS];
    const int n = ax.n;
    const int G = fft_group_lines(n);
    const int grp = blockIdx.x;
    if ((int64_t)grp >= fft_groups(AXIS, G, nx, ny, nz)) return;
    const double *__restrict__ src = FROM_SRC ? J.src : J.data;
    double *__restrict__ dst = J.data;
    const double out_scale = SCALE ? J.inv_n : 1.0;
    const double *__restrict__ lamx = J.lam[0];
    const double *__restrict__ lamy = J.lam[1];
    const double *__restrict__ lamz = J.lam[2];
    const int nlines = (G + 1) / 2;
    const int64_t estride = AXIS == 0 ? 1 : (AXIS == 1 ? (int64_t)nx : (int64_t)nx * ny);

    load_lines<AXIS>(src, buf, grp, G, n, estride, nx, ny, nz);
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
                              hipStream_t st)
{
    if (njobs <= 0 || max_groups <= 0) return hipSuccess;
    const dim3 g((unsigned)max_groups, (unsigned)njobs), b(kFftBlock);
    if (rs == 0)
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 0, FROM_SRC, SCALE>), g, b, 0, st, jobs);
    else if (rs == 1)
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 1, FROM_SRC, SCALE>), g, b, 0, st, jobs);
    else
        hipLaunchKernelGGL((k_fft_pass<KIND, AXIS, 2, FROM_SRC, SCALE>), g, b, 0, st, jobs);
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
