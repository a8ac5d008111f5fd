// ohpic_kernels.hpp -- host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ohpic_internal.hpp"

namespace ohpic {

enum ParticleMode : int {
    kModeFused = 0,       // gather + push + wrap + deposit(new x) + re-bin
    kModeDeposit = 1,     // deposit current positions
    kModeGatherPush = 2,  // gather + push in place (no wrap, no re-bin)
    kModeBoundary = 3,    // wrap + re-bin
    // as kModeFused / kModeBoundary, but every particle goes to a contiguous
    // list [0, count[par^1][0]) of buffer par^1 instead of a bin (the retry
    // after a bin overflowed)
    kModeFusedList = 4,
    kModeBoundaryList = 5,
    // as kModeFused over a resident grid: workgroups claim work items from a
    // device counter (launch_particles_queue)
    kModeFusedQueue = 6
};

struct Pcg32Dev {
    uint64_t state, inc;
};

bool tile_supported(int tx, int ty, int tz);

// parity[item.inst] selects the current buffer of each item's instance.
hipError_t launch_particles(int mode, int precision, int weight_order, const int T[3],
                            const InstDev *insts, const WorkItem *items, int nitems,
                            const int *parity, hipStream_t st);
// Bin a contiguous list [0, n) of buffer lbuf into the front runs of buffer
// lbuf^1 (count first, then scatter once capacities are set); wrap != 0
// applies the periodic wrap twice to each position first.
hipError_t launch_bin_count(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                            const int T[3], int wrap, hipStream_t st);
hipError_t launch_bin_scatter(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                              const int T[3], int wrap, hipStream_t st);
hipError_t launch_bins_to_list(const InstDev *insts, int inst, int s, int src, int dst,
                               const int *tiles, int t0, int ntiles, const int64_t *list_off,
                               hipStream_t st);
// bins tiles[i] of species s: base[i], cap[i], and used[i] particles as one
// front run of buffer par
hipError_t launch_set_bins(const InstDev *insts, int inst, int s, int par, const int *tiles,
                           const int64_t *base, const int32_t *cap, const int64_t *used, int n,
                           hipStream_t st);
hipError_t launch_cell_counts(const InstDev *insts, int inst, int s, int par, int ntiles,
                              int32_t *counts, hipStream_t st);

// The fused step as one resident grid of `blocks` workgroups that claim the
// nitems work items in list order from *queue (which must be 0 at launch:
// launch_prepare_next_b zeroes it).  particles_queue_blocks: workgroups of that kernel
// resident on the whole device.
hipError_t launch_particles_queue(int precision, int weight_order, const int T[3],
                                  const InstDev *insts, const WorkItem *items, int nitems,
                                  const int *parity, unsigned int *queue, int blocks,
                                  hipStream_t st);
int particles_queue_blocks(int precision, int weight_order, const int T[3], int device);

// mixed-precision particles (ohpic_particles_mixed.hip)
hipError_t launch_particles_mixed(int mode, int weight_order, const int T[3],
                                  const InstDev *insts, const WorkItem *items, int nitems,
                                  const int *parity, hipStream_t st);
hipError_t launch_particles_mixed_queue(int weight_order, const int T[3], const InstDev *insts,
                                        const WorkItem *items, int nitems, const int *parity,
                                        unsigned int *queue, int blocks, hipStream_t st);
int particles_mixed_queue_blocks(int weight_order, const int T[3]);
hipError_t launch_bin_count_mx(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                               const int T[3], hipStream_t st);
hipError_t launch_bin_scatter_mx(const InstDev *insts, int inst, int s, int lbuf, int64_t n,
                                 const int T[3], hipStream_t st);
hipError_t launch_bins_to_list_mx(const InstDev *insts, int inst, int s, int src, int dst,
                                  const int *tiles, int t0, int ntiles, const int64_t *list_off,
                                  hipStream_t st);
hipError_t launch_cell_counts_mx(const InstDev *insts, int inst, int s, int par, int ntiles,
                                 int32_t *counts, const int T[3], hipStream_t st);

// fields (ohpic_fields.hip)
hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v,
                           hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
hipError_t launch_fill_f64(double *p, int64_t n, double v, hipStream_t st);
// Batched grid kernels: every launch covers instances [0, ninst) of the
// table (blockIdx.y = instance).
// rho := int64 charge grid of the current buffer * 2^-charge_bits
hipError_t launch_rho_from_acc_b(const InstDev *insts, const int *parity, int ninst,
                                 int64_t max_ng, hipStream_t st);
// node E from phi + block partials of sum rho*phi, sum E^2, sum rho; with
// reduce, the partials are then reduced into diag[kDiagWe..kDiagQnet] (else
// launch_end_step_b does it)
// max_blocks: the largest efield_blocks_of() over the instances
hipError_t launch_efield_b(const InstDev *insts, int ninst, int max_blocks, bool reduce,
                           hipStream_t st);
// rho from the current charge grid and the zeroing of launch_prepare_next_b,
// in one launch at the start of a step
hipError_t launch_begin_step_b(const InstDev *insts, const int *parity, int ninst,
                               int64_t max_ng, int max_ntiles, unsigned int *queue,
                               hipStream_t st);
// launch_reduce_ke_b, the field-energy reduction of launch_efield_b, the
// tight-bin check of the new buffer (parity ^ 1) and (out not null, group
// steps) flags to out[i * kFlagSlots + k] with a parity flip unless instance i
// overflowed a bin, in one launch at the end of a step
hipError_t launch_end_step_b(const InstDev *insts, int *parity, int ninst,
                             unsigned long long *out, hipStream_t st);
// zero the next buffer's bin words and charge grid, the flags, and *queue
// (if not null: the counter of a following launch_particles_queue)
hipError_t launch_prepare_next_b(const InstDev *insts, const int *parity, int ninst,
                                 int64_t max_ng, int max_ntiles, unsigned int *queue,
                                 hipStream_t st);
// diag[kDiagKeOld..] := sums of ke_partial over the instance's items
hipError_t launch_reduce_ke_b(const InstDev *insts, int ninst, hipStream_t st);

// seeded loaders (ohpic_fields.hip, ohpic_particles_mixed.hip)
hipError_t launch_load_uniform(Pcg32Dev rng0, uint64_t offset, int64_t np, int nx, int ny,
                               int nz, double vth, double *x, double *y, double *z, double *vx,
                               double *vy, double *vz, int64_t *pid, int64_t pid0,
                               hipStream_t st);
hipError_t launch_load_uniform_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                                  uint64_t offset, int64_t np, double vth, const int T[3],
                                  hipStream_t st);
hipError_t launch_load_hump_mx(const InstDev *insts, int inst, int s, int b, Pcg32Dev rng0,
                               uint64_t offset, uint64_t budget, int64_t np, double amp,
                               double sigma, double vth, unsigned long long *fail,
                               const int T[3], hipStream_t st);
hipError_t launch_load_hump(Pcg32Dev rng0, uint64_t offset, uint64_t budget, int64_t np, int nx,
                            int ny, int nz, double amp, double sigma, double vth, double *x,
                            double *y, double *z, double *vx, double *vy, double *vz,
                            int64_t *pid, int64_t pid0, unsigned long long *fail,
                            hipStream_t st);

}  // namespace ohpic
