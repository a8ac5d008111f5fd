// ohpic_internal.hpp -- device-side data layout shared by the gfx950 kernels
// and the host runtime of libohpic.so.
//
// HBM layout of one simulation instance
//   particles : per species, two SoA buffers (current / next).  FP64 mode holds
//               x, y, z, vx, vy, vz as double; mixed mode holds cell (int32),
//               in-cell offsets ox, oy, oz and vx, vy, vz as float.  Particles
//               are binned by tile (a TX*TY*TZ block of cells); bin t of a
//               species occupies [bin_base[t], bin_base[t] + bin_cap[t]) of both
//               buffers and holds count[par][t] particles (capacity slack
//               absorbs the per-step exchange between neighbouring tiles).
//   charge    : two int64 fixed-point grids (current / next) of n nodes;
//               integer sums make the deposit independent of summation order.
//   fields    : rho, phi, Ex, Ey, Ez as double, n nodes each, x fastest.
//   a step whose exchange overflows a bin is discarded and re-run after a
//   full re-bin (repack) of the untouched input buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ohpic_bins.hpp"

namespace ohpic {

constexpr int kMaxSpecies = 8;
constexpr int kBlock = 256;  // threads per workgroup (4 wave64)

// diagnostics accumulator slots (double) per instance
enum DiagSlot : int {
    kDiagKeOld = 0,                     // [kMaxSpecies]
    kDiagKeNew = kDiagKeOld + kMaxSpecies,  // [kMaxSpecies]
    kDiagWe = kDiagKeNew + kMaxSpecies,     // sum rho*phi
    kDiagWe2 = kDiagWe + 1,                 // sum |E|^2
    kDiagQnet = kDiagWe2 + 1,               // sum rho
    kDiagSlots = kDiagQnet + 1
};

// flag slots (uint64) per instance
enum FlagSlot : int {
    kFlagOverflow = 0,  // particles that found their destination bin full
    kFlagFast = 1,      // particles whose new position left the deposit halo
    kFlagRange = 2,     // positions outside [0, n) (after the wraps when wrapping)
    kFlagTight = 3,     // a step left a bin short of free slots (bin_tight)
    kFlagListed = 4,    // entries the step wrote to tight_list (bin_listed bins)
    kFlagSlots = 5
};
// flags[kFlagEndBlocks]: blocks of k_end_step_b that finished (a counter,
// not a reported flag); the start of a step clears every word
constexpr int kFlagEndBlocks = kFlagSlots;
constexpr int kFlagWords = kFlagSlots + 1;

// tight_list entries: the bin's particle count, species and tile
constexpr int kTightListCap = 4096;
__host__ __device__ inline unsigned long long tight_entry(int s, int t, int64_t used)
{
    return (unsigned long long)used << 32 | (unsigned long long)s << 27 | (unsigned long long)t;
}
__host__ __device__ inline int tight_entry_species(unsigned long long e) { return (int)(e >> 27) & 31; }
__host__ __device__ inline int tight_entry_tile(unsigned long long e) { return (int)(e & 0x7ffffffull); }
__host__ __device__ inline int64_t tight_entry_used(unsigned long long e) { return (int64_t)(e >> 32); }

struct SpecDev {
    void *buf[2][7];       // component arrays of the two particle buffers
    int64_t *pid[2];       // optional ids
    int64_t *bin_base;     // [ntiles]
    int32_t *bin_cap;      // [ntiles]
    unsigned long long *count[2];  // [ntiles] front | back << 32 (see above)
    double qm;             // charge-to-mass ratio
    double q;              // macro charge
    double mass;           // q / qm
    double qfix;           // q * 2^charge_bits
};

struct InstDev {
    int n[3];
    int T[3];
    int nt[3];
    int ntiles;
    int nspecies;
    int weight_order;
    int track_ids;
    int precision;
    double dt;
    double b0[3];
    double fix_scale;      // 2^charge_bits
    double fix_inv;        // 2^-charge_bits
    double tight_frac;     // free-slot fraction of a tight bin (bin_tight)
    unsigned long long *tight_list;  // [kTightListCap] bins k_end_step_b found listed
    SpecDev sp[kMaxSpecies];
    unsigned long long *rho_acc[2];
    double *rho;
    double *phi;
    double *e[3];
    double *diag;                 // kDiagSlots
    double *ke_partial;           // [kKeSlots][nitems], each row reduced by one block
    double *e_partial;            // [kEfieldMaxBlocks][3] field-energy partial sums
    int nitems;                   // work items of this instance (row length of ke_partial)
    unsigned long long *flags;    // kFlagSlots
};

struct WorkItem {
    int32_t inst;
    int32_t tile;
    int16_t chunk;     // this item's slice of every bin of the tile
    int16_t nchunks;   // slices per bin
    int32_t slot;      // row of the per-item partial-sum table
};
// largest chunks per tile a configuration may ask for (WorkItem::nchunks)
constexpr int kMaxChunks = 32767;

constexpr int kKeSlots = 2 * kMaxSpecies;  // per item: ke_old[8], ke_new[8]
#ifndef OHPIC_EFIELD_MAX_BLOCKS
#define OHPIC_EFIELD_MAX_BLOCKS 4096
#endif
constexpr int kEfieldMaxBlocks = OHPIC_EFIELD_MAX_BLOCKS;  // rows of e_partial

// k_efield_b: x points per thread (2, with 16-byte accesses, for even nx)
// and x rows per 256-thread block (rows of fewer than 256 threads share one).
// An instance's block count depends on its grid only, so its energy partial
// sums do not depend on the other instances of a group.
__host__ __device__ inline int efield_vec(int nx) { return (nx & 1) ? 1 : 2; }
__host__ __device__ inline int efield_rows_per_block(int nx)
{
    const int tpr = nx / efield_vec(nx);  // threads per row
    return tpr >= 256 ? 1 : 256 / tpr;
}
__host__ __device__ inline int efield_blocks_of(int nx, int ny, int nz)
{
    const int rpb = efield_rows_per_block(nx);
    const int64_t need = ((int64_t)ny * nz + rpb - 1) / rpb;
    return (int)(need < kEfieldMaxBlocks ? need : kEfieldMaxBlocks);
}

// Records the message ohpic_last_error(NULL) returns (host side).
void set_global_error(const char *msg);

// ---------------------------------------------------------------- helpers
// Every particle and grid buffer is device (global) memory, but pointers read
// from the instance table or a job table are generic to the compiler, which
// then emits flat instructions; those also count against the LDS wait
// counter, so each wait for an LDS access would wait for every load and store
// in flight.
// Accesses through these helpers compile to global_* instructions.
template <typename T>
using gmem_t = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ gmem_t<T> *as_global(T *p)
{
    return (gmem_t<T> *)p;
}

template <typename T>
__device__ __forceinline__ T gload(const T *p, int64_t i)
{
    return as_global(p)[i];
}

template <typename T>
__device__ __forceinline__ void gstore(T *p, int64_t i, T v)
{
    as_global(p)[i] = v;
}

__device__ __forceinline__ long long fix_round(double w, double qfix)
{
    // round(w * qfix) to an integer for |w * qfix| < 2^51: one FMA against
    // 1.5 * 2^52 leaves the rounded integer in the low mantissa bits.
    const double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    double y = __fma_rn(w, qfix, kMagic);
    return __double_as_longlong(y) - 0x4338000000000000LL;
}

__host__ __device__ __forceinline__ int bin_front(unsigned long long w)
{
    return (int)(w & 0xffffffffull);
}

__host__ __device__ __forceinline__ int bin_back(unsigned long long w) { return (int)(w >> 32); }


// slot of the v-th particle of a bin with front run F and back run B
__device__ __forceinline__ int bin_slot(int v, int F, int B, int cap)
{
    return v < F ? v : cap - B + (v - F);
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear workgroup id,
// and each XCD has its own L2.  xcd_remap(b, count) turns dispatch index b
// (< count) into a logical index such that XCD k receives the contiguous
// logical range [k*per, (k+1)*per) in order (per = count / 8; the count % 8
// last ids map to themselves), so that neighbouring logical workgroups share
// an L2.  A bijection on [0, count); grids round their x size up to a multiple
// of 8 so that blockIdx.x % 8 is the XCD for every blockIdx.y.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_remap(int b, int count)
{
    const int full = count - count % kXcds;
    if (b >= full) return b;
    return (b % kXcds) * (full / kXcds) + b / kXcds;
}
__host__ __device__ inline int64_t round_up_xcds(int64_t n) { return (n + kXcds - 1) / kXcds * kXcds; }

// i mod n for i >= -n: one compare per side when i lies within a period of
// [0, n) (tile and halo indices of grids at least a tile wide), a division
// only beyond that
__device__ __forceinline__ int wrap_near(int i, int n)
{
    if (i < 0) i += n;
    if (i >= n) {
        i -= n;
        if (i >= n) i %= n;
    }
    return i;
}

__device__ __forceinline__ int wrap_index(int i, int n)
{
    // i in [-n, 2n) -> [0, n)
    if (i < 0) i += n;
    if (i >= n) i -= n;
    return i;
}

}  // namespace ohpic
