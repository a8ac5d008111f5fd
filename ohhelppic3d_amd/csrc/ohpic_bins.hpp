// ohpic_bins.hpp -- bin capacities and the planning of bin re-layouts.
//
// Plain C++ (no HIP): the host runtime of libohpic.so plans with it, the
// device kernels call the constexpr tests, and the CPU tests compile it alone.
//
// Each species keeps one bin per tile in both particle buffers: bin t occupies
// [base[t], base[t] + cap[t]).  A full re-bin lays the bins out in tile order
// from element 0; bins that later run short of free slots move one by one
// into the pool beyond them (the rest of the buffer's alloc elements).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace ohpic {

// A bin holds fewer than 2^29 slots: the particle kernels address a bin with
// 32-bit byte offsets.
constexpr int64_t kBinCapLimit = (int64_t)1 << 29;

// Capacity of a bin holding c particles: the slack fraction and 16 slots on
// top, rounded up to 64.
inline int64_t bin_capacity(int64_t c, double slack)
{
    const int64_t v = (int64_t)std::ceil((double)c * (1.0 + slack)) + 16;
    return (v + 63) / 64 * 64;
}

// A bin is tight when fewer than max(kTightMinFree, frac * cap) of its slots
// are free; a bin whose capacity bin_capacity just set is not.
constexpr int kTightMinFree = 16;
constexpr bool bin_tight(int64_t cap, int64_t used, double frac)
{
    return cap - used < kTightMinFree || (double)(cap - used) < frac * (double)cap;
}

// Free-slot fraction of a tight bin: half the slack a bin is given.
inline double tight_fraction(double slack) { return slack / (2.0 * (1.0 + slack)); }

// A bin moves along with the tight ones when fewer than max(kTightMinFree,
// 1.5 frac * cap) of its slots are free (three quarters of its slack used):
// in a region whose density grows, the bins next in line move in the same
// pass.  Every tight bin is listed; a bin whose capacity bin_capacity just set
// (with the slack, or twice the slack for a moved bin) is not.
constexpr bool bin_listed(int64_t cap, int64_t used, double frac)
{
    return cap - used < kTightMinFree || (double)(cap - used) < 1.5 * frac * (double)cap;
}

// Relocation pool: each buffer holds this fraction of the slack (times the
// species' capacity) beyond the bins set_bins lays out, where tight bins move.
constexpr double kPoolSlackShare = 0.5;

// A bin to move: its tile and the particles it holds.
struct BinCand {
    int tile;
    int64_t used;
};

// A bin moved into the pool.
struct BinMove {
    int tile;
    int64_t base;
    int32_t cap;
    int64_t used;
};

enum PlanResult : int { kPlanMoves = 1, kPlanWhole = 2, kPlanTooBig = 3 };

// Plans the relocation of a species' listed bins (cands, in any order).
// Every live bin lies below *pool; each listed bin moves to the pool with
// capacity bin_capacity(used, 2 * slack) (a bin that filled up is likely in a
// region that keeps growing; at most kBinCapLimit - 64), packed one after
// another, and *pool advances past them.  A moved bin's old slots stay unused
// until the next full re-bin.
// Returns kPlanMoves (moves and *pool set), kPlanWhole (the pool up to alloc
// cannot take them: re-bin the species in full) or kPlanTooBig (bin *big_bin
// needs kBinCapLimit slots or more).
inline int plan_relocation(const std::vector<BinCand> &cands, double slack, int64_t *pool,
                           int64_t alloc, std::vector<BinMove> &moves, int *big_bin)
{
    moves.clear();
    int64_t p = *pool;
    for (const BinCand &c : cands) {
        if (bin_capacity(c.used, slack) >= kBinCapLimit) {
            *big_bin = c.tile;
            moves.clear();
            return kPlanTooBig;
        }
        const int64_t k = std::min(bin_capacity(c.used, 2.0 * slack), kBinCapLimit - 64);
        moves.push_back({c.tile, p, (int32_t)k, c.used});
        p += k;
    }
    if (p > alloc) {
        moves.clear();
        return kPlanWhole;
    }
    *pool = p;
    return kPlanMoves;
}

}  // namespace ohpic
