/*
 * ohpic_oracle.c -- CPU FP64 restatement of ohhelppic3d's per-timestep loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the reported CPU baseline).  The
 * product library never links or calls this code.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/particle_mover/particle_mover_boris.f90:21-64       orc_boris_move
 *   src/interpolator/linear_interpolator.f90:22-69          orc_interp / orc_gather_periodic
 *   src/scatter/linear_scatter.f90:22-55                    orc_scatter / orc_deposit_periodic
 *   src/particle_boundary/periodic_particle_boundary.f90:42-53   orc_apply_periodic
 *   src/particle_boundary/particle_boundaries.f90:23-43     (x, y, z order in orc_step)
 *   src/field_solver/poisson/poisson_solver3d.f90:104-132,134-207   orc_poisson_periodic
 *   src/field_solver/fft/fftw3/mpi_fftw3_solver.f90:151-173,209-236  orc_r2r_axis (R2HC/HC2R)
 *   src/field_solver/poisson_field_solver.f90:90-175        orc_efield
 *   src/ohhelppic3d.f90:286-381                             orc_step (step order)
 *   src/utils/science_constants.f90:4                       ORC_REF_PI
 *   lib/pcg-c/include/pcg_variants.h:156-159,279,588-597,790-798,1375-1380
 *   lib/pcg-c/src/pcg-advance-64.c:44-59, src/pcg32_double.c:60-63   PCG32
 *   src/particle_injector/.../position_random_uniform_distribution.f90:31-42
 *   src/particle_injector/.../velocity_maxwell_distribution3d.f90:33-58 (loader)
 *
 * Resolutions of undefined reference behaviour implemented here:
 *   mirrored CIC weights kept as weight_order 0; standard order is 1;
 *   E_y/E_z read component 1 of phi (the reference indexes a 1-component
 *   field with 2/3);  node E = (E_edge(i-e) + E_edge(i))/2 with periodic wrap;
 *   deposited charge is folded periodically (node n == node 0);
 *   B is a uniform constant b0 (the reference never sets it);
 *   the Maxwellian loader uses std = vth and log(1-u) (PCG can return 0);
 *   the boundary is applied twice per particle as mainstep does.
 */
#include "ohpic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_REF_PI 3.14159265358979 /* science_constants.f90:4 (truncated) */
#define ORC_TWO_PI_EXACT 6.283185307179586476925286766559

/* ======================================================================== */
/* PCG32                                                                     */
/* ======================================================================== */
#define ORC_PCG_MULT 6364136223846793005ULL /* pcg_variants.h:279 */

static inline void orc_pcg_step(orc_pcg32 *rng)
{
    /* pcg_variants.h:588-591 */
    rng->state = rng->state * ORC_PCG_MULT + rng->inc;
}

void orc_pcg32_srandom(orc_pcg32 *rng, uint64_t initstate, uint64_t initseq)
{
    /* pcg_variants.h:790-798 */
    rng->state = 0u;
    rng->inc = (initseq << 1u) | 1u;
    orc_pcg_step(rng);
    rng->state += initstate;
    orc_pcg_step(rng);
}

uint32_t orc_pcg32_random(orc_pcg32 *rng)
{
    /* pcg_variants.h:1375-1380 with the XSH-RR output of :156-159 */
    uint64_t old = rng->state;
    orc_pcg_step(rng);
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

uint32_t orc_pcg32_boundedrand(orc_pcg32 *rng, uint32_t bound)
{
    /* Lemire-free rejection form used by the pcg-c boundedrand_r functions. */
    uint32_t threshold = (-bound) % bound;
    for (;;) {
        uint32_t r = orc_pcg32_random(rng);
        if (r >= threshold)
            return r % bound;
    }
}

double orc_pcg32_double(orc_pcg32 *rng)
{
    /* pcg32_double.c:60-63: ldexp(u32, -32) */
    return ldexp((double)orc_pcg32_random(rng), -32);
}

void orc_pcg32_advance(orc_pcg32 *rng, uint64_t delta)
{
    /* pcg-advance-64.c:44-59 (Brown's O(log n) LCG jump) */
    uint64_t cur_mult = ORC_PCG_MULT, cur_plus = rng->inc;
    uint64_t acc_mult = 1u, acc_plus = 0u;
    while (delta > 0) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    rng->state = acc_mult * rng->state + acc_plus;
}

/* ======================================================================== */
/* Loaders                                                                   */
/* ======================================================================== */
static void orc_maxwell3(orc_pcg32 *rng, double vth, double *v1, double *v2, double *v3)
{
    /* velocity_maxwell_distribution3d.f90:33-58 draw order (4 uniforms), with
     * std = vth (not 2*vth) and log(1-u) so u == 0 is safe; no 6-sigma clip. */
    double r1 = orc_pcg32_double(rng);
    double r2 = orc_pcg32_double(rng);
    double r3 = orc_pcg32_double(rng);
    double r4 = orc_pcg32_double(rng);
    double a = vth * sqrt(-2.0 * log(1.0 - r1));
    double b = vth * sqrt(-2.0 * log(1.0 - r3));
    *v1 = a * cos(2.0 * ORC_REF_PI * r2);
    *v2 = a * sin(2.0 * ORC_REF_PI * r2);
    *v3 = b * sin(2.0 * ORC_REF_PI * r4);
}

int64_t orc_load_uniform_maxwell(orc_pcg32 *rng, int64_t np, int nx, int ny, int nz,
                                 double vth, double *x, double *y, double *z,
                                 double *vx, double *vy, double *vz)
{
    /* particle_injector_with_distribution.f90:51: random_fix(n) consumes one
     * draw before the particle loop. */
    (void)orc_pcg32_double(rng);
    for (int64_t i = 0; i < np; ++i) {
        /* position_random_uniform_distribution.f90:41: rand*(upper-lower)+lower */
        x[i] = orc_pcg32_double(rng) * (double)nx + 0.0;
        y[i] = orc_pcg32_double(rng) * (double)ny + 0.0;
        z[i] = orc_pcg32_double(rng) * (double)nz + 0.0;
        orc_maxwell3(rng, vth, &vx[i], &vy[i], &vz[i]);
    }
    return 1 + 7 * np;
}

int orc_load_hump_maxwell(const orc_pcg32 *rng0, uint64_t base, uint64_t budget,
                          int64_t np, int nx, int ny, int nz, double amp, double sigma,
                          double vth, double *x, double *y, double *z,
                          double *vx, double *vy, double *vz)
{
    const double cx = 0.5 * nx, cy = 0.5 * ny, cz = 0.5 * nz;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    const uint64_t max_cand = (budget - 4u) / 4u;
    for (int64_t i = 0; i < np; ++i) {
        orc_pcg32 rng = *rng0;
        orc_pcg32_advance(&rng, base + (uint64_t)i * budget);
        uint64_t c;
        for (c = 0; c < max_cand; ++c) {
            double px = orc_pcg32_double(&rng) * (double)nx;
            double py = orc_pcg32_double(&rng) * (double)ny;
            double pz = orc_pcg32_double(&rng) * (double)nz;
            double u = orc_pcg32_double(&rng);
            double dx = px - cx, dy = py - cy, dz = pz - cz;
            double r2 = dx * dx + dy * dy + dz * dz;
            double f = (1.0 + amp * exp(-r2 * inv2s2)) / (1.0 + amp);
            if (u < f) {
                x[i] = px;
                y[i] = py;
                z[i] = pz;
                break;
            }
        }
        if (c == max_cand)
            return -1;
        orc_maxwell3(&rng, vth, &vx[i], &vy[i], &vz[i]);
    }
    return 0;
}

/* ======================================================================== */
/* Per-particle operators                                                    */
/* ======================================================================== */
double orc_apply_periodic(int n, double p)
{
    /* periodic_particle_boundary.f90:46-52: the two comparisons in this order */
    if (p < 0)
        p = (double)n + p;
    if (p >= n)
        p = p - (double)n;
    return p;
}

static inline void orc_cross(const double a[3], const double b[3], double c[3])
{
    /* futils m_vector%cross (not vendored): the right-handed cross product */
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void orc_boris_move(double *x, double *y, double *z, double *vx, double *vy, double *vz,
                    double qm, const double eb[6], double dt)
{
    /* particle_mover_boris.f90:36-62 */
    double dt2 = dt * 0.5;
    double t[3], s[3], upm[3], upa[3], upp[3], c[3];
    for (int i = 0; i < 3; ++i)
        t[i] = qm * eb[3 + i] * dt2;
    double tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    for (int i = 0; i < 3; ++i)
        s[i] = 2 * t[i] / (1 + tt);
    double uold[3] = {*vx, *vy, *vz};
    for (int i = 0; i < 3; ++i)
        upm[i] = uold[i] + qm * eb[i] * dt2;
    orc_cross(upm, t, c);
    for (int i = 0; i < 3; ++i)
        upa[i] = upm[i] + c[i];
    orc_cross(upa, s, c);
    for (int i = 0; i < 3; ++i)
        upp[i] = upm[i] + c[i];
    *vx = upp[0] + qm * eb[0] * dt2;
    *vy = upp[1] + qm * eb[1] * dt2;
    *vz = upp[2] + qm * eb[2] * dt2;
    *x = *x + *vx * dt;
    *y = *y + *vy * dt;
    *z = *z + *vz * dt;
}

static inline void orc_weights(double r, int weight_order, double w[2])
{
    /* linear_interpolator.f90:54-57 / linear_scatter.f90:41-42: the lower
     * node (i=1) gets r and the upper node 1-r (weight_order 0). */
    if (weight_order == 0) {
        w[0] = r;
        w[1] = 1.0 - r;
    } else {
        w[0] = 1.0 - r;
        w[1] = r;
    }
}

static inline int64_t orc_field_index(int c, int nelem, const int lb[3], const int ub[3],
                                      int ix, int iy, int iz)
{
    int64_t sx = ub[0] - lb[0] + 1, sy = ub[1] - lb[1] + 1;
    return c + (int64_t)nelem * ((ix - lb[0]) + sx * ((iy - lb[1]) + sy * (int64_t)(iz - lb[2])));
}

void orc_interp(const double *values, int nelem, const int lb[3], const int ub[3],
                const double pos[3], int weight_order, double *ret)
{
    /* linear_interpolator.f90:22-69 (subdomain origin 0: to_local_position is
     * the identity) */
    int i0[3];
    double wx[2], wy[2], wz[2];
    for (int a = 0; a < 3; ++a)
        i0[a] = (int)pos[a];
    orc_weights(pos[0] - i0[0], weight_order, wx);
    orc_weights(pos[1] - i0[1], weight_order, wy);
    orc_weights(pos[2] - i0[2], weight_order, wz);
    for (int c = 0; c < nelem; ++c)
        ret[c] = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int c = 0; c < nelem; ++c) {
                    double v = values[orc_field_index(c, nelem, lb, ub, i0[0] + i, i0[1] + j,
                                                      i0[2] + k)];
                    ret[c] = ret[c] + v * wx[i] * wy[j] * wz[k];
                }
}

void orc_scatter(double *values, int nelem, const int lb[3], const int ub[3],
                 const double pos[3], const double *amount, int weight_order)
{
    /* linear_scatter.f90:22-55 */
    int i0[3];
    double wx[2], wy[2], wz[2];
    for (int a = 0; a < 3; ++a)
        i0[a] = (int)pos[a];
    orc_weights(pos[0] - i0[0], weight_order, wx);
    orc_weights(pos[1] - i0[1], weight_order, wy);
    orc_weights(pos[2] - i0[2], weight_order, wz);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int c = 0; c < nelem; ++c) {
                    int64_t idx = orc_field_index(c, nelem, lb, ub, i0[0] + i, i0[1] + j,
                                                  i0[2] + k);
                    values[idx] = values[idx] + amount[c] * wx[i] * wy[j] * wz[k];
                }
}

/* ======================================================================== */
/* Periodic whole-grid operators                                             */
/* ======================================================================== */
void orc_deposit_periodic(int nx, int ny, int nz, int64_t np, const double *x,
                          const double *y, const double *z, double q, int weight_order,
                          double *rho)
{
    /* mainstep deposit loop (ohhelppic3d.f90:311-320) with linear_scatter's
     * formula; node index taken mod n (periodic fold of the ghost node). */
    for (int64_t p = 0; p < np; ++p) {
        int ix = (int)x[p], iy = (int)y[p], iz = (int)z[p];
        double wx[2], wy[2], wz[2];
        orc_weights(x[p] - ix, weight_order, wx);
        orc_weights(y[p] - iy, weight_order, wy);
        orc_weights(z[p] - iz, weight_order, wz);
        for (int i = 0; i < 2; ++i) {
            int gx = (ix + i) % nx;
            for (int j = 0; j < 2; ++j) {
                int gy = (iy + j) % ny;
                for (int k = 0; k < 2; ++k) {
                    int gz = (iz + k) % nz;
                    int64_t idx = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
                    rho[idx] = rho[idx] + q * wx[i] * wy[j] * wz[k];
                }
            }
        }
    }
}

void orc_gather_periodic(int nx, int ny, int nz, const double *ex, const double *ey,
                         const double *ez, double px, double py, double pz,
                         int weight_order, double out[3])
{
    int ix = (int)px, iy = (int)py, iz = (int)pz;
    double wx[2], wy[2], wz[2];
    orc_weights(px - ix, weight_order, wx);
    orc_weights(py - iy, weight_order, wy);
    orc_weights(pz - iz, weight_order, wz);
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < 2; ++i) {
        int gx = (ix + i) % nx;
        for (int j = 0; j < 2; ++j) {
            int gy = (iy + j) % ny;
            for (int k = 0; k < 2; ++k) {
                int gz = (iz + k) % nz;
                int64_t idx = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
                out[0] = out[0] + ex[idx] * wx[i] * wy[j] * wz[k];
                out[1] = out[1] + ey[idx] * wx[i] * wy[j] * wz[k];
                out[2] = out[2] + ez[idx] * wx[i] * wy[j] * wz[k];
            }
        }
    }
}

/* ---- complex DFT (recursive mixed radix, any n) --------------------------- */
typedef struct {
    double re, im;
} orc_cpx;

static inline orc_cpx orc_cis(int64_t m, int64_t n, int sign)
{
    /* exp(sign * 2 pi i m / n) with m reduced into [0, n) */
    m %= n;
    if (m < 0)
        m += n;
    double ang = ORC_TWO_PI_EXACT * (double)m / (double)n;
    orc_cpx w = {cos(ang), sign * sin(ang)};
    return w;
}

static int orc_smallest_factor(int n)
{
    if (n % 4 == 0)
        return 4;
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0)
            return p;
    return n;
}

/* out[k] = sum_j in[j*s] exp(sign 2 pi i j k / n) */
static void orc_dft(int n, int s, const orc_cpx *in, orc_cpx *out, int sign, orc_cpx *tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = orc_smallest_factor(n);
    if (p == n) {
        for (int k = 0; k < n; ++k) {
            double re = 0, im = 0;
            for (int j = 0; j < n; ++j) {
                orc_cpx w = orc_cis((int64_t)j * k, n, sign);
                orc_cpx a = in[(int64_t)j * s];
                re += a.re * w.re - a.im * w.im;
                im += a.re * w.im + a.im * w.re;
            }
            out[k].re = re;
            out[k].im = im;
        }
        return;
    }
    int m = n / p;
    for (int q = 0; q < p; ++q)
        orc_dft(m, s * p, in + (int64_t)q * s, tmp + (int64_t)q * m, sign, out);
    /* combine: X[k + m u] = sum_q (w_n^{q k} Y_q[k]) w_p^{q u} */
    orc_cpx t[64];
    orc_cpx *tt = p <= 64 ? t : (orc_cpx *)malloc(sizeof(orc_cpx) * p);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            orc_cpx w = orc_cis((int64_t)q * k, n, sign);
            orc_cpx y = tmp[(int64_t)q * m + k];
            tt[q].re = y.re * w.re - y.im * w.im;
            tt[q].im = y.re * w.im + y.im * w.re;
        }
        for (int u = 0; u < p; ++u) {
            double re = 0, im = 0;
            for (int q = 0; q < p; ++q) {
                orc_cpx w = orc_cis((int64_t)q * u, p, sign);
                re += tt[q].re * w.re - tt[q].im * w.im;
                im += tt[q].re * w.im + tt[q].im * w.re;
            }
            out[k + (int64_t)m * u].re = re;
            out[k + (int64_t)m * u].im = im;
        }
    }
    if (tt != t)
        free(tt);
}

void orc_r2hc_1d(int n, const double *in, double *out)
{
    /* FFTW R2HC layout (test_mpi_fftw3_solver.f90:78):
     * r0, r1, ..., r_{n/2}, i_{(n+1)/2-1}, ..., i1, forward sign -1 */
    orc_cpx *a = (orc_cpx *)calloc(3 * (size_t)n, sizeof(orc_cpx));
    orc_cpx *b = a + n, *t = a + 2 * n;
    for (int j = 0; j < n; ++j) {
        a[j].re = in[j];
        a[j].im = 0.0;
    }
    orc_dft(n, 1, a, b, -1, t);
    for (int k = 0; k <= n / 2; ++k)
        out[k] = b[k].re;
    for (int k = 1; k < (n + 1) / 2; ++k)
        out[n - k] = b[k].im;
    free(a);
}

void orc_hc2r_1d(int n, const double *in, double *out)
{
    /* FFTW HC2R: unnormalised inverse of R2HC (HC2R(R2HC(x)) = n x) */
    orc_cpx *a = (orc_cpx *)calloc(3 * (size_t)n, sizeof(orc_cpx));
    orc_cpx *b = a + n, *t = a + 2 * n;
    a[0].re = in[0];
    a[0].im = 0.0;
    for (int k = 1; k < (n + 1) / 2; ++k) {
        a[k].re = in[k];
        a[k].im = in[n - k];
        a[n - k].re = in[k];
        a[n - k].im = -in[n - k];
    }
    if (n % 2 == 0) {
        a[n / 2].re = in[n / 2];
        a[n / 2].im = 0.0;
    }
    orc_dft(n, 1, a, b, +1, t);
    for (int j = 0; j < n; ++j)
        out[j] = b[j].re;
    free(a);
}

static int64_t orc_line_count(int nx, int ny, int nz, int axis)
{
    return axis == 0 ? (int64_t)ny * nz : (axis == 1 ? (int64_t)nx * nz : (int64_t)nx * ny);
}

/* Transforms the lines [l0, l1) of one axis; a line's first element is
 * nx*l (x lines), (l % nx) + nx*ny*(l / nx) (y lines) or l (z lines). */
static void orc_r2r_lines(int nx, int ny, int nz, int axis, int kind, double *arr,
                          int64_t l0, int64_t l1)
{
    int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    int64_t stride = axis == 0 ? 1 : (axis == 1 ? nx : (int64_t)nx * ny);
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *res = line + n;
    for (int64_t l = l0; l < l1; ++l) {
        int64_t base = axis == 0 ? (int64_t)nx * l
                     : (axis == 1 ? l % nx + (int64_t)nx * ny * (l / nx) : l);
        for (int j = 0; j < n; ++j)
            line[j] = arr[base + j * stride];
        if (kind == 0)
            orc_r2hc_1d(n, line, res);
        else
            orc_hc2r_1d(n, line, res);
        for (int j = 0; j < n; ++j)
            arr[base + j * stride] = res[j];
    }
    free(line);
}

void orc_r2r_axis(int nx, int ny, int nz, int axis, int kind, double *arr)
{
    orc_r2r_lines(nx, ny, nz, axis, kind, arr, 0, orc_line_count(nx, ny, nz, axis));
}

double orc_wave_number_sq(int k, int n)
{
    /* poisson_solver3d.f90:111-117 (periodic branch), squared as at :72;
     * n is the periodic length (global_block%sizes - 1). */
    double wn;
    if (k <= n / 2)
        wn = 2.0 * sin(ORC_REF_PI * k / (double)n);
    else
        wn = 2.0 * sin(ORC_REF_PI * (n - k) / (double)n);
    return wn * wn;
}

void orc_poisson_periodic(int nx, int ny, int nz, const double *rho, double *phi)
{
    /* poisson_solver3d.f90:134-207: phi := rho; forward R2HC (x, y, z);
     * divide by sum of squared modified wave numbers; DC := 0; HC2R; / N */
    int64_t n = (int64_t)nx * ny * nz;
    memcpy(phi, rho, sizeof(double) * (size_t)n);
    orc_r2r_axis(nx, ny, nz, 0, 0, phi);
    orc_r2r_axis(nx, ny, nz, 1, 0, phi);
    orc_r2r_axis(nx, ny, nz, 2, 0, phi);
    for (int kz = 0; kz < nz; ++kz) {
        double wz = orc_wave_number_sq(kz, nz);
        for (int ky = 0; ky < ny; ++ky) {
            double wy = orc_wave_number_sq(ky, ny);
            for (int kx = 0; kx < nx; ++kx) {
                double wx = orc_wave_number_sq(kx, nx);
                int64_t idx = kx + (int64_t)nx * (ky + (int64_t)ny * kz);
                if (kx == 0 && ky == 0 && kz == 0)
                    phi[idx] = 0.0;
                else
                    phi[idx] = phi[idx] / (wx + wy + wz);
            }
        }
    }
    orc_r2r_axis(nx, ny, nz, 0, 1, phi);
    orc_r2r_axis(nx, ny, nz, 1, 1, phi);
    orc_r2r_axis(nx, ny, nz, 2, 1, phi);
    double norm = (double)nx * (double)ny * (double)nz; /* product(normalizers) */
    for (int64_t i = 0; i < n; ++i)
        phi[i] = phi[i] / norm;
}

static void orc_efield_planes(int nx, int ny, int nz, const double *phi, double *ex,
                              double *ey, double *ez, int z0, int z1)
{
    /* poisson_field_solver.f90:139-144 edge difference E_a(i) = phi(i)-phi(i+e_a)
     * and :165-170 node average E_a(i) = 0.5*(E_a(i-e_a) + E_a(i)), periodic */
    for (int iz = z0; iz < z1; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                int xp = (ix + 1) % nx, xm = (ix + nx - 1) % nx;
                int yp = (iy + 1) % ny, ym = (iy + ny - 1) % ny;
                int zp = (iz + 1) % nz, zm = (iz + nz - 1) % nz;
#define ORC_IDX(a, b, c) ((a) + (int64_t)nx * ((b) + (int64_t)ny * (c)))
                double p0 = phi[ORC_IDX(ix, iy, iz)];
                double edge_x = p0 - phi[ORC_IDX(xp, iy, iz)];
                double edge_xm = phi[ORC_IDX(xm, iy, iz)] - p0;
                double edge_y = p0 - phi[ORC_IDX(ix, yp, iz)];
                double edge_ym = phi[ORC_IDX(ix, ym, iz)] - p0;
                double edge_z = p0 - phi[ORC_IDX(ix, iy, zp)];
                double edge_zm = phi[ORC_IDX(ix, iy, zm)] - p0;
                int64_t o = ORC_IDX(ix, iy, iz);
#undef ORC_IDX
                ex[o] = 0.5 * (edge_xm + edge_x);
                ey[o] = 0.5 * (edge_ym + edge_y);
                ez[o] = 0.5 * (edge_zm + edge_z);
            }
}

void orc_efield(int nx, int ny, int nz, const double *phi, double *ex, double *ey,
                double *ez)
{
    orc_efield_planes(nx, ny, nz, phi, ex, ey, ez, 0, nz);
}

/* ======================================================================== */
/* One timestep                                                              */
/* ======================================================================== */
double orc_kinetic_energy(int64_t np, const double *vx, const double *vy, const double *vz,
                          double mass)
{
    double s = 0.0;
    for (int64_t i = 0; i < np; ++i)
        s += vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
    return 0.5 * mass * s;
}

static void orc_push_range(orc_sim *s, int sp, int64_t p0, int64_t p1)
{
    /* gather -> boris_move -> boundary applied twice (ohhelppic3d.f90:335-358) */
    const double qm = s->qm[sp];
    double *x = s->x[sp], *y = s->y[sp], *z = s->z[sp];
    double *vx = s->vx[sp], *vy = s->vy[sp], *vz = s->vz[sp];
    for (int64_t p = p0; p < p1; ++p) {
        double eb[6];
        orc_gather_periodic(s->nx, s->ny, s->nz, s->ex, s->ey, s->ez, x[p], y[p], z[p],
                            s->weight_order, eb);
        eb[3] = s->b0[0];
        eb[4] = s->b0[1];
        eb[5] = s->b0[2];
        orc_boris_move(&x[p], &y[p], &z[p], &vx[p], &vy[p], &vz[p], qm, eb, s->dt);
        for (int pass = 0; pass < 2; ++pass) {
            x[p] = orc_apply_periodic(s->nx, x[p]);
            y[p] = orc_apply_periodic(s->ny, y[p]);
            z[p] = orc_apply_periodic(s->nz, z[p]);
        }
    }
}

void orc_step(orc_sim *s)
{
    /* mainstep order (ohhelppic3d.f90:311-358): rho := 0, deposit every
     * species; field solve; gather, Boris, boundary (applied twice). */
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const int64_t ng = (int64_t)nx * ny * nz;
    memset(s->rho, 0, sizeof(double) * (size_t)ng);
    for (int sp = 0; sp < s->nspecies; ++sp)
        orc_deposit_periodic(nx, ny, nz, s->np[sp], s->x[sp], s->y[sp], s->z[sp], s->q[sp],
                             s->weight_order, s->rho);
    orc_poisson_periodic(nx, ny, nz, s->rho, s->phi);
    orc_efield(nx, ny, nz, s->phi, s->ex, s->ey, s->ez);
    double we = 0.0, qnet = 0.0;
    for (int64_t i = 0; i < ng; ++i) {
        we += s->rho[i] * s->phi[i];
        qnet += s->rho[i];
    }
    s->field_energy = 0.5 * we;
    s->net_charge = qnet;
    for (int sp = 0; sp < s->nspecies; ++sp) {
        const double mass = s->q[sp] / s->qm[sp];
        s->ke_old[sp] = orc_kinetic_energy(s->np[sp], s->vx[sp], s->vy[sp], s->vz[sp], mass);
        orc_push_range(s, sp, 0, s->np[sp]);
        s->ke_new[sp] = orc_kinetic_energy(s->np[sp], s->vx[sp], s->vy[sp], s->vz[sp], mass);
    }
}

/* ---- the same step on several host threads -------------------------------- */
/* The CPU baseline's multi-threaded form: every particle and every FFT line
 * goes through the same functions as in orc_step, split over threads; the
 * deposit goes to one charge array per thread, summed in thread order.  Only
 * rho (and the sums) can differ from orc_step, by summation order. */
#ifdef _OPENMP
#include <omp.h>
#define ORC_TID() omp_get_thread_num()
#define ORC_NT() omp_get_num_threads()
#else
#define ORC_TID() 0
#define ORC_NT() 1
#endif

static void orc_r2r_axis_mt(int nx, int ny, int nz, int axis, int kind, double *arr, int nt)
{
    const int64_t nl = orc_line_count(nx, ny, nz, axis);
#pragma omp parallel num_threads(nt)
    {
        const int64_t t = ORC_TID(), T = ORC_NT();
        orc_r2r_lines(nx, ny, nz, axis, kind, arr, nl * t / T, nl * (t + 1) / T);
    }
}

int orc_step_threads(orc_sim *s, int nthreads)
{
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    const int nt = 1;
    (void)nthreads;
#endif
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const int64_t ng = (int64_t)nx * ny * nz;
    double *part = (double *)malloc(sizeof(double) * (size_t)ng * (size_t)nt);
    if (!part)
        return -1;
    int used = 1;
#pragma omp parallel num_threads(nt)
    {
        const int t = ORC_TID(), T = ORC_NT();
        if (t == 0)
            used = T;
        double *mine = part + (size_t)ng * (size_t)t;
        memset(mine, 0, sizeof(double) * (size_t)ng);
        for (int sp = 0; sp < s->nspecies; ++sp) {
            const int64_t n = s->np[sp], p0 = n * t / T, p1 = n * (t + 1) / T;
            orc_deposit_periodic(nx, ny, nz, p1 - p0, s->x[sp] + p0, s->y[sp] + p0,
                                 s->z[sp] + p0, s->q[sp], s->weight_order, mine);
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < ng; ++i) {
            double r = 0.0;
            for (int k = 0; k < T; ++k)
                r += part[(size_t)ng * (size_t)k + (size_t)i];
            s->rho[i] = r;
        }
    }
    free(part);

    /* field solve: orc_poisson_periodic with its loops split over threads */
    double *phi = s->phi;
    memcpy(phi, s->rho, sizeof(double) * (size_t)ng);
    for (int a = 0; a < 3; ++a)
        orc_r2r_axis_mt(nx, ny, nz, a, 0, phi, used);
#pragma omp parallel for schedule(static) num_threads(used)
    for (int kz = 0; kz < nz; ++kz) {
        double wz = orc_wave_number_sq(kz, nz);
        for (int ky = 0; ky < ny; ++ky) {
            double wy = orc_wave_number_sq(ky, ny);
            for (int kx = 0; kx < nx; ++kx) {
                double wx = orc_wave_number_sq(kx, nx);
                int64_t idx = kx + (int64_t)nx * (ky + (int64_t)ny * kz);
                if (kx == 0 && ky == 0 && kz == 0)
                    phi[idx] = 0.0;
                else
                    phi[idx] = phi[idx] / (wx + wy + wz);
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        orc_r2r_axis_mt(nx, ny, nz, a, 1, phi, used);
    const double norm = (double)nx * (double)ny * (double)nz;
    double we = 0.0, qnet = 0.0;
#pragma omp parallel num_threads(used)
    {
#pragma omp for schedule(static)
        for (int64_t i = 0; i < ng; ++i)
            phi[i] = phi[i] / norm;
        const int t = ORC_TID(), T = ORC_NT();
        orc_efield_planes(nx, ny, nz, phi, s->ex, s->ey, s->ez, (int)((int64_t)nz * t / T),
                          (int)((int64_t)nz * (t + 1) / T));
#pragma omp for schedule(static) reduction(+ : we, qnet)
        for (int64_t i = 0; i < ng; ++i) {
            we += s->rho[i] * phi[i];
            qnet += s->rho[i];
        }
    }
    s->field_energy = 0.5 * we;
    s->net_charge = qnet;

    for (int sp = 0; sp < s->nspecies; ++sp) {
        const double mass = s->q[sp] / s->qm[sp];
        const int64_t n = s->np[sp];
        const double *vx = s->vx[sp], *vy = s->vy[sp], *vz = s->vz[sp];
        double ko = 0.0, kn = 0.0;
#pragma omp parallel num_threads(used)
        {
#pragma omp for schedule(static) reduction(+ : ko)
            for (int64_t p = 0; p < n; ++p)
                ko += vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
            const int64_t t = ORC_TID(), T = ORC_NT();
            orc_push_range(s, sp, n * t / T, n * (t + 1) / T);
#pragma omp barrier
#pragma omp for schedule(static) reduction(+ : kn)
            for (int64_t p = 0; p < n; ++p)
                kn += vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
        }
        s->ke_old[sp] = 0.5 * mass * ko;
        s->ke_new[sp] = 0.5 * mass * kn;
    }
    return used;
}

void orc_cell_counts(int nx, int ny, int nz, int64_t np, const double *x, const double *y,
                     const double *z, int32_t *counts)
{
    /* int() truncation of the position, linear id ix + nx*(iy + ny*iz) */
    for (int64_t p = 0; p < np; ++p) {
        int ix = (int)x[p], iy = (int)y[p], iz = (int)z[p];
        counts[ix + (int64_t)nx * (iy + (int64_t)ny * iz)] += 1;
    }
}
