// ohpic_bins.hpp -- bin capacities and the planning of bin re-layouts.
//
// Plain C++ (no HIP): the host runtime of libohpic.so plans with it, the
// device kernels call the constexpr tests, and the CPU tests compile it alone.
//
// Each species keeps one bin per tile, laid out in tile order in both particle
// buffers: bin t occupies [base[t], base[t] + cap[t]), base[t] + cap[t] <=
// base[t + 1], the last bin ending at or before the buffer's alloc elements.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
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

// A re-layout of the bins a..b (inclusive).
struct BinWindow {
    int a, b;
};

enum PlanResult : int { kPlanNone = 0, kPlanWindows = 1, kPlanWhole = 2, kPlanTooBig = 3 };

// Plans the re-layout of a species' tight bins.  Each tight bin and its
// neighbours in tile order form a window, grown (to the right, then to the
// left, alternately) until the capacities bin_capacity gives their used
// counts fit into the window's slots, from the first bin's base to the next
// window-external base (the last bin also owns the buffer's tail).  Windows
// stay disjoint: one that can grow neither way merges with the previous one.
// The window's bins get those capacities plus the window's spare slots shared
// out in 64-slot units (each bin below kBinCapLimit).
// Returns kPlanNone (no tight bin), kPlanWindows (win, nbase and ncap set;
// nbase/ncap equal base/cap outside the windows), kPlanWhole (only the whole
// buffer fits: re-bin the species in full) or kPlanTooBig (bin *big_bin needs
// kBinCapLimit slots or more).
inline int plan_rebalance(int nt, const std::vector<int64_t> &base,
                          const std::vector<int32_t> &cap, int64_t alloc,
                          const std::vector<int64_t> &used, double tight_frac, double slack,
                          std::vector<BinWindow> &win, std::vector<int64_t> &nbase,
                          std::vector<int32_t> &ncap, int *big_bin)
{
    win.clear();
    std::vector<int64_t> pre((size_t)nt + 1, 0);
    bool any = false;
    for (int t = 0; t < nt; ++t) {
        any = any || bin_tight(cap[t], used[t], tight_frac);
        pre[t + 1] = pre[t] + bin_capacity(used[t], slack);
    }
    if (!any) return kPlanNone;
    auto end_of = [&](int b) { return b + 1 < nt ? base[b + 1] : alloc; };
    auto fits = [&](int a, int b) { return pre[b + 1] - pre[a] <= end_of(b) - base[a]; };
    for (int t = 0; t < nt; ++t) {
        if (!bin_tight(cap[t], used[t], tight_frac) || (!win.empty() && t <= win.back().b))
            continue;
        int a = t, b = t;
        int lo = win.empty() ? 0 : win.back().b + 1;
        while (!fits(a, b)) {
            const bool right = b + 1 < nt, left = a > lo;
            if (!right && !left) {
                if (win.empty()) return kPlanWhole;
                a = win.back().a;
                win.pop_back();
                lo = win.empty() ? 0 : win.back().b + 1;
                continue;
            }
            if (right) ++b;
            if (left && !fits(a, b)) --a;
        }
        win.push_back({a, b});
    }
    nbase.assign(base.begin(), base.end());
    ncap.assign(cap.begin(), cap.end());
    for (const BinWindow &w : win) {
        const int nw = w.b - w.a + 1;
        const int64_t units = (end_of(w.b) - base[w.a] - (pre[w.b + 1] - pre[w.a])) / 64;
        int64_t o = base[w.a];
        for (int t = w.a; t <= w.b; ++t) {
            const int64_t need = pre[t + 1] - pre[t];
            if (need >= kBinCapLimit) {
                *big_bin = t;
                return kPlanTooBig;
            }
            const int64_t share = units / nw + ((t - w.a) < units % nw ? 1 : 0);
            const int64_t c = std::min<int64_t>(need + 64 * share, kBinCapLimit - 64);
            nbase[t] = o;
            ncap[t] = (int32_t)c;
            o += c;
        }
    }
    return kPlanWindows;
}

}  // namespace ohpic
