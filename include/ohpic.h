/*
 * ohpic.h -- C ABI of the MI355X (gfx950) implementation of ohhelppic3d's
 * per-timestep loop: CIC charge scatter, periodic spectral Poisson solve,
 * node-centred E, CIC field gather + Boris push, periodic particle boundary,
 * and the per-step re-binning that replaces OhHelp's particle migration.
 *
 * Plain C: pointers, sizes and status codes only.  Every call returns
 * OHPIC_OK (0) or a negative OHPIC_E* code; ohpic_last_error(ctx) gives the
 * message (ctx may be NULL for errors raised before a context exists).  The
 * reference aborts on error (error stop / MPI_Abort); a Fortran caller turns a
 * non-zero status into `error stop msg`.
 *
 * Ownership: the context owns every device allocation; host buffers passed in
 * are owned by the caller and only read or written during the call.  One
 * context = one simulation instance = one HIP stream; a context is not
 * thread-safe, separate contexts are independent.
 *
 * Units and layout follow the reference: dx = 1, eps0 = 1, periodic grid of
 * nx*ny*nz nodes, field arrays node-centred with x fastest
 * (values(:, ix, iy, iz) order); particle positions in [0, n) per axis.
 */
#ifndef OHPIC_H
#define OHPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OHPIC_MAX_SPECIES 8

/* status codes */
#define OHPIC_OK 0
#define OHPIC_EINVAL -1   /* bad argument */
#define OHPIC_EHIP -2     /* HIP runtime error */
#define OHPIC_ENOMEM -3   /* device memory / bin capacity exhausted */
#define OHPIC_ESTATE -4   /* call not valid in the current state */
#define OHPIC_ENODEV -6   /* no usable GPU */
#define OHPIC_EIO -7      /* file or HDF5 error */
#define OHPIC_ENOLIB -8   /* a library loaded at run time (HDF5) is missing */

/* precision modes */
#define OHPIC_FP64 0      /* FP64 positions/velocities and arithmetic */
#define OHPIC_MIXED 1     /* int32 cell + FP32 offset, FP32 v; FP64 fields */

/* CIC weight order (linear_interpolator.f90:54-57, linear_scatter.f90:41-42) */
#define OHPIC_WEIGHTS_REFERENCE 0 /* lower node gets r = frac(x) (as written) */
#define OHPIC_WEIGHTS_STANDARD 1  /* lower node gets 1 - r */

/* field selectors for get/put */
#define OHPIC_FIELD_RHO 0
#define OHPIC_FIELD_PHI 1
#define OHPIC_FIELD_EX 2
#define OHPIC_FIELD_EY 3
#define OHPIC_FIELD_EZ 4
#define OHPIC_FIELD_EB 5  /* 6 components: Ex, Ey, Ez, B0x, B0y, B0z */

typedef struct ohpic_ctx ohpic_ctx;
typedef struct ohpic_group ohpic_group;

typedef struct {
    int32_t nx, ny, nz;
    int32_t nspecies;
    double qm[OHPIC_MAX_SPECIES];        /* charge_to_mass_ratio (parameters.f90:127) */
    double q_macro[OHPIC_MAX_SPECIES];   /* charge_per_macro_particle (parameters.f90:140-149) */
    int64_t capacity[OHPIC_MAX_SPECIES]; /* max particles per species */
    double dt;
    double b0[3];          /* uniform magnetic field */
    int32_t precision;     /* OHPIC_FP64 or OHPIC_MIXED */
    int32_t weight_order;  /* OHPIC_WEIGHTS_* */
    int32_t tile[3];       /* cells per tile; 0 = default (8^3 if the grid holds
                              at least 512 such tiles, else 4^3) */
    double slack;          /* bin capacity slack fraction; <= 0 = 0.15 */
    int32_t chunks;        /* work items per tile, 0 to 32767; 0 = fill the device */
    int32_t track_ids;     /* keep a 64-bit particle id beside each particle */
    int32_t device;        /* HIP device ordinal */
    int32_t charge_bits;   /* fixed-point fraction bits of the charge grid; 0 = derived */
} ohpic_config;

typedef struct {
    int64_t step;                          /* completed steps */
    double ke_old[OHPIC_MAX_SPECIES];      /* sum 1/2 m |v^{n-1/2}|^2 of the last push */
    double ke_new[OHPIC_MAX_SPECIES];      /* sum 1/2 m |v^{n+1/2}|^2 of the last push */
    double field_energy;                   /* 1/2 sum rho*phi of the last solve */
    double field_energy_e2;                /* 1/2 sum |E|^2 of the last solve */
    double net_charge;                     /* sum rho of the last solve */
    double total_energy;                   /* (ke_old+ke_new)/2 summed + field_energy */
    int64_t repacks;                       /* full re-bins so far */
    int64_t fast_particles;                /* particles that moved > 1 cell in a step */
    int64_t rebalances;                    /* passes that moved bins short of slack */
    int64_t compactions;                   /* compactions of every species' bins (a
                                              relocation pool ran out) */
} ohpic_diag;

/* --- context --------------------------------------------------------------- */
int ohpic_config_default(ohpic_config *cfg);
int ohpic_create(const ohpic_config *cfg, ohpic_ctx **out);
int ohpic_destroy(ohpic_ctx *ctx);
const char *ohpic_last_error(const ohpic_ctx *ctx);
int ohpic_get_config(const ohpic_ctx *ctx, ohpic_config *out);

/* --- particles (oh_particle storage, ohparticles.f90:16-78) -----------------
 * SoA host arrays; positions must lie in [0, n).  pid may be NULL. */
int ohpic_load_particles(ohpic_ctx *ctx, int ispec, int64_t n, const double *x,
                         const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz, const int64_t *pid);
int ohpic_particle_count(ohpic_ctx *ctx, int ispec, int64_t *n);
/* Copies the species' particles out in bin order (pid gives the identity). */
int ohpic_store_particles(ohpic_ctx *ctx, int ispec, int64_t cap, int64_t *n, double *x,
                          double *y, double *z, double *vx, double *vy, double *vz,
                          int64_t *pid);
/* AoS interchange in the oh_particle record layout: x,y,z,vx,vy,vz (f64),
 * pid (i64), nid, spec (i32; spec is 1-based as in the reference). */
typedef struct {
    double x, y, z, vx, vy, vz;
    int64_t pid;
    int32_t nid, spec;
} ohpic_oh_particle;
int ohpic_load_oh_particles(ohpic_ctx *ctx, int64_t n, const ohpic_oh_particle *rec);
int ohpic_store_oh_particles(ohpic_ctx *ctx, int64_t cap, int64_t *n, ohpic_oh_particle *rec);

/* Seeded loaders generated on the device.  Draw order follows the reference
 * injector (particle_injector_with_distribution.f90:51-72): one shared PCG32
 * stream (pcg_generator.f90:21-41), species in order, one random_fix draw per
 * species, then per particle x, y, z (u*n) and four Box-Muller uniforms with
 * std vth[s].  np[s] particles per species. */
int ohpic_seed_uniform_maxwell(ohpic_ctx *ctx, uint64_t initstate, uint64_t initseq,
                               const int64_t *np, const double *vth);
/* Density 1 + amp*exp(-|r-c|^2/(2 sigma^2)) by rejection sampling; particle i
 * of species s draws from the sub-stream at offset (s*2^40 + i*budget). */
int ohpic_seed_hump_maxwell(ohpic_ctx *ctx, uint64_t initstate, uint64_t initseq,
                            const int64_t *np, const double *vth, double amp, double sigma,
                            uint64_t budget);

/* --- the per-step operators --------------------------------------------------
 * deposit      : rho := sum_s q_s W(x)   (linear_scatter.f90:22-55 over all
 *                particles, ohhelppic3d.f90:311-320)
 * field_solve  : rho -> phi -> E       (poisson_field_solver.f90:90-175)
 * gather_push  : E(x) gather + Boris  (linear_interpolator.f90:22-69,
 *                particle_mover_boris.f90:21-64; ohhelppic3d.f90:347-349)
 * boundary     : periodic wrap (applied twice) + re-bin
 *                (periodic_particle_boundary.f90:42-53; ohhelppic3d.f90:351-355)
 * step         : nsteps of deposit -> solve -> gather_push -> boundary, fused. */
int ohpic_deposit(ohpic_ctx *ctx);
int ohpic_field_solve(ohpic_ctx *ctx);
int ohpic_gather_push(ohpic_ctx *ctx);
int ohpic_boundary(ohpic_ctx *ctx);
int ohpic_step(ohpic_ctx *ctx, int nsteps);
int ohpic_synchronize(ohpic_ctx *ctx);

/* --- fields (t_OhField, ohfield.f90:53-72) -----------------------------------
 * Copy between the device grid and a host array values(nelem, lb:ub) in
 * Fortran order; component `comp` (0-based) of the host array receives the
 * field (for OHPIC_FIELD_EB all 6 components, comp ignored).  Halo nodes
 * outside [0, n) are filled periodically on get; put reads [0, n) only. */
int ohpic_get_field(ohpic_ctx *ctx, int which, double *values, int nelem, const int32_t lb[3],
                    const int32_t ub[3], int comp);
int ohpic_put_field(ohpic_ctx *ctx, int which, const double *values, int nelem,
                    const int32_t lb[3], const int32_t ub[3], int comp);

/* --- FFT (t_MPIFFTSolver3d forward/backward, mpi_fftw3_solver.f90:209-236) --
 * Separable 3-D real-to-real transform of an nx*ny*nz host array (x fastest):
 * forward = R2HC on every axis (FFTW halfcomplex layout); backward = HC2R on
 * every axis divided by nx*ny*nz.  in and out may alias. */
int ohpic_fft3d_r2r(int device, int nx, int ny, int nz, int backward, const double *in,
                    double *out);

/* Allocation-free ohpic_fft3d_r2r.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results bit for bit equal to
 * ohpic_fft3d_r2r's; it calls no hipMalloc/hipFree and creates no stream (it
 * runs on the null stream).  It works in one fixed device array of
 * OHPIC_FFT_NOALLOC_MAX_POINTS doubles plus the axis tables (about 16 MiB),
 * which each GPU it runs on holds until the process ends; nx*ny*nz above
 * that maximum returns OHPIC_EINVAL.  Every call on a GPU shares that array:
 * the library serialises its calls with a process-wide lock, and the array
 * keeps the last call's contents. */
#define OHPIC_FFT_NOALLOC_MAX_POINTS (1 << 21)
int ohpic_fft3d_r2r_noalloc(int device, int nx, int ny, int nz, int backward, const double *in,
                            double *out);

/* --- per-particle operators over host arrays -----------------------------------
 * The reference's per-particle bindings, batched: each call uploads the
 * arrays, runs one kernel on `device` and copies the results back.
 * ohpic_boris_batch   : t_ParticleMoverBoris%move (particle_mover_boris.f90:21-64)
 *                       eb holds 6 values per particle (Ex,Ey,Ez,Bx,By,Bz)
 * ohpic_interp_batch  : t_LinearInterpolator%interp (linear_interpolator.f90:22-69)
 *                       pos holds x,y,z per particle; values(nelem, lb:ub) in
 *                       Fortran order; out holds nelem values per particle
 * ohpic_scatter_batch : t_LinearScatter%scatter (linear_scatter.f90:22-55), adds
 *                       amount(nelem) * W(x) into values(nelem, lb:ub)
 * ohpic_wrap_batch    : apply_periodic (periodic_particle_boundary.f90:42-53)
 *                       applied `applications` times to each p[i] (axis length nn)
 * A stencil reaching outside lb:ub is an error (the reference reads out of
 * bounds there). */
int ohpic_boris_batch(int device, int64_t n, double *x, double *y, double *z, double *vx,
                      double *vy, double *vz, const double *eb, double qm, double dt);
int ohpic_interp_batch(int device, int64_t n, const double *pos, const double *values, int nelem,
                       const int32_t lb[3], const int32_t ub[3], int weight_order, double *out);
int ohpic_scatter_batch(int device, int64_t n, const double *pos, double *values, int nelem,
                        const int32_t lb[3], const int32_t ub[3], const double *amount,
                        int weight_order);
int ohpic_wrap_batch(int device, int64_t n, double *p, int32_t nn, int applications);

/* --- diagnostics ------------------------------------------------------------*/
int ohpic_diagnostics(ohpic_ctx *ctx, ohpic_diag *out);
/* The energy history of one step in four numbers: out[0] kinetic energy
 * (time-centred, all species), out[1] field energy 1/2 sum rho*phi, out[2]
 * their sum, out[3] net charge sum rho (as ohpic_diagnostics reports them). */
int ohpic_energy(ohpic_ctx *ctx, double out[4]);
/* per-cell particle counts (cell ix + nx*(iy + ny*iz)); ispec < 0 = all */
int ohpic_cell_counts(ohpic_ctx *ctx, int ispec, int32_t *counts);
/* Kernel timing (HIP events on the context stream).  enable != 0 starts
 * recording; ohpic_kernel_times fills up to n (name, total_ms, launches). */
int ohpic_enable_timing(ohpic_ctx *ctx, int enable);
int ohpic_kernel_times(ohpic_ctx *ctx, int n, char names[][32], double *total_ms,
                       int64_t *launches, int *count);
/* Device memory held by the context, bytes. */
int ohpic_memory_footprint(ohpic_ctx *ctx, int64_t *bytes);
/* Predicted device bytes for a configuration (used for admission). */
int ohpic_estimate_footprint(const ohpic_config *cfg, int64_t *bytes);

/* --- groups: many independent instances advanced together ------------------
 * No counterpart in the reference (it runs one simulation per MPI job).  A
 * group holds contexts of one device, precision, weight order and tile shape
 * and advances all of them with one launch per kernel and step: the particle
 * kernel walks a work list of (instance, tile, chunk) items of every member,
 * the FFT passes a table of per-instance transforms of different sizes.  The
 * per-step launch sequence is captured once in a hipGraph and replayed until
 * membership or a member's bins change.  Each member's result is the same as
 * when it is stepped alone with ohpic_step (charge deposit is integer, so the
 * fields and particles match bit for bit; the kinetic-energy sums may differ
 * in the last bits because particle order inside a bin is not fixed).
 * Members stay usable through the single-context calls between group steps.
 * Group errors: ohpic_group_last_error(g). */
int ohpic_group_create(int device, ohpic_group **out);
int ohpic_group_destroy(ohpic_group *g); /* detaches every member */
int ohpic_group_attach(ohpic_group *g, ohpic_ctx *ctx);
int ohpic_group_detach(ohpic_group *g, ohpic_ctx *ctx);
int ohpic_group_size(const ohpic_group *g, int *n);
int ohpic_group_contains(const ohpic_group *g, const ohpic_ctx *ctx, int *yes);
/* Advance every member by nsteps (same sequence as ohpic_step), in launch
 * sequences of `batch` steps (ohpic_group_set_batch; default 8) while that
 * many remain, then single steps; each sequence is one graph launch and one
 * host synchronisation.  A member whose bins overflow in a sequence keeps its
 * state for the rest of it; afterwards that step is redone as a list and the
 * member's remaining steps are taken alone, so every member ends nsteps
 * further.  When a member's step fails (e.g. a particle moved more than two
 * box lengths) the other members still complete the sequence, the failed
 * members are detached (their contexts keep the message and refuse further
 * steps) and the first failure's code is returned; ohpic_group_stats counts
 * the steps taken. */
int ohpic_group_step(ohpic_group *g, int nsteps);
/* steps per launch sequence, 1 to 16 */
int ohpic_group_set_batch(ohpic_group *g, int steps);
/* enable != 0 (default): replay captured hipGraphs; 0: launch directly */
int ohpic_group_use_graph(ohpic_group *g, int enable);
/* enable != 0 (default): the particle step of all members is one persistent
 * launch (a resident grid whose workgroups claim work items, heaviest first,
 * from a device counter); 0: one workgroup per work item */
int ohpic_group_use_persistent(ohpic_group *g, int enable);
/* steps taken, graph captures, work items of the current table */
int ohpic_group_stats(const ohpic_group *g, int64_t *steps, int64_t *captures, int64_t *items);
/* launch sequences replayed from a captured graph, and sequences enqueued
 * launch by launch (graphs disabled, or another thread was inside a library
 * call when the sequence needed a capture: a capture is only taken while no
 * other thread makes HIP calls through the library, since such calls can
 * invalidate it) */
int ohpic_group_sequences(const ohpic_group *g, int64_t *replayed, int64_t *direct);
const char *ohpic_group_last_error(const ohpic_group *g);

/* --- autotuner ---------------------------------------------------------------
 * No counterpart in the reference.  Times the group step over candidate tile
 * shapes (4x4x4, 8x8x8, 8x8x4) and chunks per tile (1, 2, 4) on up to `sample`
 * of the n configurations (evenly spaced), each seeded with capacity[s]
 * Maxwellian particles, and returns the fastest (tile, chunks) with its
 * measured particle-steps/s.  The configurations' own tile/chunks are ignored. */
typedef struct {
    int32_t tile[3];
    int32_t chunks;
    double particle_steps_per_s;
    int32_t candidates; /* candidates timed */
} ohpic_tune_result;
int ohpic_autotune(const ohpic_config *cfgs, int n, int sample, int steps,
                   ohpic_tune_result *out);

/* --- snapshots (t_Hdf5ForOhfield, src/hdf5/hdf5_for_ohfield.f90:11-89) -------
 * One HDF5 file holding one group, with one 3-D FP64 dataset per write, as
 * the reference's phi output 'phisp00_0000.h5' / group 'phisp' / datasets '1',
 * '2', ... (ohhelppic3d.f90:79-96).  Shapes and offsets are in Fortran order
 * (x first) as the reference passes them to the Fortran HDF5 API; through the
 * C API the dataset dimensions read (z, y, x) with x fastest.  HDF5 >= 1.10 is
 * loaded at run time (OHPIC_HDF5_LIB, else libhdf5.so.103 / libhdf5.so, then
 * /opt/conda/lib): without it these return OHPIC_ENOLIB.  Calls are serialised
 * process-wide. */
typedef struct ohpic_snapshot ohpic_snapshot;
/* new_Hdf5ForOhfield(filename, group_name, global_shape, global_offset, comm)
 * (hdf5_for_ohfield.f90:27-43; new_HDF5File 'w' + create_group, hdf5.f90:46-89):
 * creates (truncates) the file and the group.  global_offset may be NULL. */
int ohpic_snapshot_open(const char *filename, const char *group, const int64_t global_shape[3],
                        const int64_t global_offset[3], ohpic_snapshot **out);
/* hdf5ForOhfield_write(dataset_name, ohfield, ps) (hdf5_for_ohfield.f90:45-76):
 * field `which` (RHO, PHI, EX, EY or EZ) of ctx as the block of nodes 0..n per
 * axis at offset 0, node n holding node 0 (periodic); the global shape must be
 * at least (nx+1, ny+1, nz+1). */
int ohpic_snapshot_write(ohpic_snapshot *s, const char *dataset, ohpic_ctx *ctx, int which);
/* write_dataset3d_double (hdf5.f90:140-185): a host block values(local_shape)
 * in Fortran order written at local_offset (NULL: 0) of a new dataset of the
 * global shape. */
int ohpic_snapshot_write_host(ohpic_snapshot *s, const char *dataset, const double *values,
                              const int64_t local_shape[3], const int64_t local_offset[3]);
/* hdf5ForOhfield_close (hdf5_for_ohfield.f90:78-86); s may be NULL. */
int ohpic_snapshot_close(ohpic_snapshot *s);
const char *ohpic_snapshot_last_error(const ohpic_snapshot *s);
/* Reads group/dataset of a snapshot file: shape[3] (Fortran order) always,
 * and the values (x fastest) when values is not NULL and capacity (in
 * doubles) holds them.  No counterpart in the reference's output path (its
 * reader, hdf5.f90:91-138, reads 1-D integer datasets). */
int ohpic_snapshot_read(const char *filename, const char *group, const char *dataset,
                        double *values, int64_t capacity, int64_t shape[3]);

#ifdef __cplusplus
}
#endif
#endif /* OHPIC_H */
