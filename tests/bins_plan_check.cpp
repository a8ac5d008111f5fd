// Property check of ohpic::plan_rebalance (ohpic_bins.hpp) on random bin
// layouts: run by tests/test_bins_plan.py.  Exit status 0 when every case
// holds; otherwise prints the failing case.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ohpic_bins.hpp"

using namespace ohpic;

static int failures = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("case %d: ", icase);           \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
            ++failures;                                \
            return;                                    \
        }                                              \
    } while (0)

static void one_case(int icase, std::mt19937_64 &rng, int counts[4])
{
    std::uniform_int_distribution<int> nt_d(1, 300);
    const int nt = nt_d(rng);
    const double slack = std::uniform_real_distribution<double>(0.01, 0.5)(rng);
    const double frac = tight_fraction(slack);
    std::vector<int64_t> base(nt), used(nt);
    std::vector<int32_t> cap(nt);
    std::uniform_int_distribution<int> c_d(0, rng() % 2 ? 200 : 20000);
    int64_t o = 0;
    for (int t = 0; t < nt; ++t) {
        const int64_t c0 = c_d(rng);
        cap[t] = (int32_t)bin_capacity(c0, slack);
        base[t] = o;
        o += cap[t];
        if (rng() % 7 == 0) o += 64 * (int64_t)(rng() % 4);  // a gap after the bin
        // the counts drift: most bins a little, some up to their capacity
        const double g = std::uniform_real_distribution<double>(-0.3, 0.3)(rng);
        int64_t u = (int64_t)(c0 * (1.0 + g));
        if (rng() % 5 == 0) u = cap[t] - (int64_t)(rng() % 20);
        used[t] = std::max<int64_t>(0, std::min<int64_t>(u, cap[t]));
    }
    const int64_t alloc = o + (int64_t)(rng() % 3) * 64 * (int64_t)(rng() % 50);
    std::vector<BinWindow> win;
    std::vector<int64_t> nbase;
    std::vector<int32_t> ncap;
    int big = -1;
    const int r = plan_rebalance(nt, base, cap, alloc, used, frac, slack, win, nbase, ncap, &big);
    counts[r] += 1;
    bool any_tight = false;
    int64_t need = 0;
    for (int t = 0; t < nt; ++t) {
        any_tight = any_tight || bin_tight(cap[t], used[t], frac);
        need += bin_capacity(used[t], slack);
    }
    if (r == kPlanNone) {
        CHECK(!any_tight, "kPlanNone with a tight bin");
        return;
    }
    CHECK(any_tight, "plan without a tight bin");
    if (r == kPlanWhole) {
        CHECK(need > alloc - base[0], "kPlanWhole although the whole buffer fits");
        return;
    }
    CHECK(r == kPlanWindows, "unexpected result %d", r);
    std::vector<char> in(nt, 0);
    int prev_b = -1;
    for (const BinWindow &w : win) {
        CHECK(w.a > prev_b && w.a <= w.b && w.b < nt, "windows out of order or overlapping");
        prev_b = w.b;
        bool has_tight = false;
        CHECK(nbase[w.a] == base[w.a], "window start moved");
        const int64_t end = w.b + 1 < nt ? base[w.b + 1] : alloc;
        for (int t = w.a; t <= w.b; ++t) {
            in[t] = 1;
            has_tight = has_tight || bin_tight(cap[t], used[t], frac);
            CHECK(ncap[t] % 64 == 0 && ncap[t] >= bin_capacity(used[t], slack) &&
                      ncap[t] < kBinCapLimit,
                  "bin %d: capacity %d for %lld particles", t, ncap[t], (long long)used[t]);
            CHECK(!bin_tight(ncap[t], used[t], frac), "bin %d still tight", t);
            if (t > w.a)
                CHECK(nbase[t] == nbase[t - 1] + ncap[t - 1], "bin %d not contiguous", t);
        }
        CHECK(has_tight, "window without a tight bin");
        CHECK(nbase[w.b] + ncap[w.b] <= end, "window overruns its span");
    }
    for (int t = 0; t < nt; ++t) {
        if (bin_tight(cap[t], used[t], frac)) CHECK(in[t], "tight bin %d not in a window", t);
        if (!in[t]) CHECK(nbase[t] == base[t] && ncap[t] == cap[t], "bin %d outside moved", t);
    }
}

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 20000;
    std::mt19937_64 rng(12345);
    int counts[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) one_case(i, rng, counts);
    std::printf("none %d windows %d whole %d toobig %d failures %d\n", counts[0], counts[1],
                counts[2], counts[3], failures);
    // bin_capacity leaves a fresh bin untight at every count
    for (int64_t c = 0; c < 200000; c += 7)
        for (double s : {0.01, 0.05, 0.15, 0.6})
            if (bin_tight(bin_capacity(c, s), c, tight_fraction(s))) {
                std::printf("fresh bin of %lld at slack %g is tight\n", (long long)c, s);
                return 1;
            }
    return failures ? 1 : 0;
}
