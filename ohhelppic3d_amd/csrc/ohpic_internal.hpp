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
//   overflow  : a list of particles whose destination bin was full; a full
//               re-bin (repack) drains it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
    kFlagOverflow = 0,  // particles appended to the overflow list (slot allocator)
    kFlagFast = 1,      // particles whose new position left the deposit halo
    kFlagLost = 2,      // overflow list full: particles dropped (fatal)
    kFlagRange = 3,     // positions still outside [0, n) after two wraps (fatal)
    kFlagSlots = 4
};

struct SpecDev {
    void *buf[2][7];       // component arrays of the two particle buffers
    int64_t *pid[2];       // optional ids
    int64_t *bin_base;     // [ntiles]
    int32_t *bin_cap;      // [ntiles]
    int32_t *count[2];     // [ntiles] particles per bin (may exceed cap: overflow)
    double qm;             // charge-to-mass ratio
    double q;              // macro charge
    double mass;           // q / qm
    double qfix;           // q * 2^charge_bits
};

struct OvfDev {
    void *comp[7];         // same component types as the particle buffers
    int64_t *pid;
    int32_t *dest;         // destination tile
    int32_t *spec;         // species
    int64_t cap;
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
    SpecDev sp[kMaxSpecies];
    unsigned long long *rho_acc[2];
    double *rho;
    double *phi;
    double *e[3];
    double *diag;                 // kDiagSlots
    unsigned long long *flags;    // kFlagSlots
    OvfDev ovf;
};

struct WorkItem {
    int32_t inst;
    int32_t tile;
    int32_t chunk;
    int32_t nchunks;
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ long long fix_round(double w, double qfix)
{
    // round(w * qfix) to an integer for |w * qfix| < 2^51: one FMA against
    // 1.5 * 2^52 leaves the rounded integer in the low mantissa bits.
    const double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    double y = __fma_rn(w, qfix, kMagic);
    return __double_as_longlong(y) - 0x4338000000000000LL;
}

__device__ __forceinline__ int wrap_index(int i, int n)
{
    // i in [-n, 2n) -> [0, n)
    if (i < 0) i += n;
    if (i >= n) i -= n;
    return i;
}

}  // namespace ohpic
