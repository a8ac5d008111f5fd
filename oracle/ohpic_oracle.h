/*
 * ohpic_oracle.h -- CPU FP64 restatement of ohhelppic3d's per-timestep loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (libohpic.so) never links it.
 *
 * Conventions
 *   - grid arrays are periodic, node-centred, no halo, x fastest:
 *       a[ix + nx*(iy + ny*iz)]   (Fortran order of t_OhField%values(:,ix,iy,iz))
 *   - particles are SoA FP64 arrays x, y, z, vx, vy, vz in grid units (dx = 1)
 *   - weight_order 0 = the reference's mirrored CIC weights (lower node gets r),
 *     weight_order 1 = standard CIC (lower node gets 1-r)
 *
 * Every function cites the reference file:line it restates.  Where the
 * reference is undefined (out-of-bounds reads, unset B, unfolded ghost charge)
 * the documented resolution is implemented instead; each resolution is named
 * at the function.
 */
#ifndef OHPIC_ORACLE_H
#define OHPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PCG32 (lib/pcg-c: pcg_setseq_64 state, XSH-RR output) --------------- */
typedef struct {
    uint64_t state;
    uint64_t inc;
} orc_pcg32;

void     orc_pcg32_srandom(orc_pcg32 *rng, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_random(orc_pcg32 *rng);
uint32_t orc_pcg32_boundedrand(orc_pcg32 *rng, uint32_t bound);
double   orc_pcg32_double(orc_pcg32 *rng);
void     orc_pcg32_advance(orc_pcg32 *rng, uint64_t delta);

/* ---- seeded loaders --------------------------------------------------------*/
/* Uniform positions + isotropic Maxwellian (std = vth) for one species, drawn
 * from `rng` in the reference injector's order.  Returns draws consumed. */
int64_t orc_load_uniform_maxwell(orc_pcg32 *rng, int64_t np, int nx, int ny, int nz,
                                 double vth, double *x, double *y, double *z,
                                 double *vx, double *vy, double *vz);

/* Gaussian-hump density n(r) ~ 1 + amp*exp(-|r-c|^2/(2 sigma^2)) by rejection
 * sampling; particle i draws from its own sub-stream at offset base + i*budget. */
int orc_load_hump_maxwell(const orc_pcg32 *rng0, uint64_t base, uint64_t budget,
                          int64_t np, int nx, int ny, int nz, double amp, double sigma,
                          double vth, double *x, double *y, double *z,
                          double *vx, double *vy, double *vz);

/* ---- per-particle operators (reference granularity) ----------------------- */
double orc_apply_periodic(int n, double p);

void orc_boris_move(double *x, double *y, double *z, double *vx, double *vy, double *vz,
                    double qm, const double eb[6], double dt);

/* Gather nelem components from a t_OhField-shaped array values(nelem, lb:ub, ...)
 * at global position pos (subdomain origin 0). */
void orc_interp(const double *values, int nelem, const int lb[3], const int ub[3],
                const double pos[3], int weight_order, double *ret);

void orc_scatter(double *values, int nelem, const int lb[3], const int ub[3],
                 const double pos[3], const double *amount, int weight_order);

/* ---- periodic whole-grid operators ---------------------------------------- */
void orc_deposit_periodic(int nx, int ny, int nz, int64_t np, const double *x,
                          const double *y, const double *z, double q, int weight_order,
                          double *rho);

void orc_gather_periodic(int nx, int ny, int nz, const double *ex, const double *ey,
                         const double *ez, double px, double py, double pz,
                         int weight_order, double out[3]);

/* FFTW-semantics real-to-real transforms along one axis of a 3-D array
 * (axis 0 = x).  kind 0 = R2HC (forward), 1 = HC2R (unnormalised backward). */
void orc_r2r_axis(int nx, int ny, int nz, int axis, int kind, double *a);
void orc_r2hc_1d(int n, const double *in, double *out);
void orc_hc2r_1d(int n, const double *in, double *out);

/* Periodic Green's-function denominator per axis, halfcomplex position k. */
double orc_wave_number_sq(int k, int n);

void orc_poisson_periodic(int nx, int ny, int nz, const double *rho, double *phi);

void orc_efield(int nx, int ny, int nz, const double *phi, double *ex, double *ey,
                double *ez);

/* ---- one full timestep ---------------------------------------------------- */
typedef struct {
    int nx, ny, nz;
    int nspecies;
    double dt;
    double b0[3];
    int weight_order;
    double qm[8];
    double q[8];
    int64_t np[8];
    double *x[8], *y[8], *z[8], *vx[8], *vy[8], *vz[8];
    double *rho, *phi, *ex, *ey, *ez;
    /* diagnostics of the last step */
    double ke_old[8]; /* sum 0.5 m |v^{n-1/2}|^2 per species */
    double ke_new[8]; /* sum 0.5 m |v^{n+1/2}|^2 per species */
    double field_energy; /* 0.5 sum rho*phi */
    double net_charge;   /* sum rho */
} orc_sim;

void orc_step(orc_sim *s);

/* orc_step with its particle and FFT-line loops split over nthreads OpenMP
 * threads (0: the OpenMP default); the deposit sums one charge array per
 * thread in thread order.  Returns the thread count used, -1 if out of
 * memory.  One thread when built without OpenMP. */
int orc_step_threads(orc_sim *s, int nthreads);

/* Energy helpers. */
double orc_kinetic_energy(int64_t np, const double *vx, const double *vy, const double *vz,
                          double mass);

/* Per-cell particle counts (cell id ix + nx*(iy + ny*iz), int() truncation). */
void orc_cell_counts(int nx, int ny, int nz, int64_t np, const double *x, const double *y,
                     const double *z, int32_t *counts);

#ifdef __cplusplus
}
#endif
#endif
