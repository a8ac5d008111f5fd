// Property check of the bin relocation planning (ohpic_bins.hpp) on random
// bin layouts: run by tests/test_bins_plan.py.  Exit status 0 when every case
// holds; otherwise prints the failing case.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ohpic_bins.hpp"

using namespace ohpic;

static int failures = 0;
#define CHECK(c, ...)                        \
    do {                                     \
        if (!(c)) {                          \
            std::printf("case %d: ", icase); \
            std::printf(__VA_ARGS__);        \
            std::printf("\n");               \
            ++failures;                      \
            return;                          \
        }                                    \
    } while (0)

static void one_case(int icase, std::mt19937_64 &rng, int counts[4])
{
    const int nt = std::uniform_int_distribution<int>(1, 300)(rng);
    const double slack = std::uniform_real_distribution<double>(0.01, 0.5)(rng);
    const double frac = tight_fraction(slack);
    std::vector<int64_t> used(nt);
    std::vector<int32_t> cap(nt);
    std::uniform_int_distribution<int> c_d(0, rng() % 2 ? 200 : 20000);
    int64_t o = 0, total = 0;
    for (int t = 0; t < nt; ++t) {
        const int64_t c0 = c_d(rng);
        total += c0;
        cap[t] = (int32_t)bin_capacity(c0, slack);
        o += cap[t];
        // the counts drift: most bins a little, some up to their capacity
        const double g = std::uniform_real_distribution<double>(-0.3, 0.1)(rng);
        int64_t u = (int64_t)(c0 * (1.0 + g));
        if (rng() % 40 == 0) u = cap[t] - (int64_t)(rng() % 20);
        used[t] = std::max<int64_t>(0, std::min<int64_t>(u, cap[t]));
    }
    // the pool of a buffer of this capacity, sometimes half in use already
    const int64_t alloc = o + (int64_t)std::ceil(total * slack * kPoolSlackShare) +
                          (int64_t)(rng() % 3) * 64;
    const int64_t pool0 = o + (rng() % 3 == 0 ? (alloc - o) / 2 : 0);
    // the candidates k_end_step_b lists
    std::vector<BinCand> cands;
    int64_t need = 0;
    for (int t = 0; t < nt; ++t) {
        if (bin_tight(cap[t], used[t], frac))
            CHECK(bin_listed(cap[t], used[t], frac), "tight bin %d not listed", t);
        if (bin_listed(cap[t], used[t], frac)) {
            cands.push_back({t, used[t]});
            need += bin_capacity(used[t], 2.0 * slack);
        }
    }
    if (cands.empty()) {
        counts[0] += 1;
        return;
    }
    int64_t pool = pool0;
    std::vector<BinMove> moves;
    int big = -1;
    const int r = plan_relocation(cands, slack, &pool, alloc, moves, &big);
    counts[r] += 1;
    if (r == kPlanWhole) {
        CHECK(pool0 + need > alloc && moves.empty() && pool == pool0,
              "kPlanWhole although the pool holds the listed bins");
        return;
    }
    CHECK(r == kPlanMoves, "unexpected result %d", r);
    CHECK(moves.size() == cands.size(), "%zu moves for %zu bins", moves.size(), cands.size());
    int64_t p = pool0;
    for (size_t k = 0; k < moves.size(); ++k) {
        const BinMove &m = moves[k];
        CHECK(m.tile == cands[k].tile && m.used == cands[k].used, "move %zu is not bin %d", k,
              cands[k].tile);
        CHECK(m.base == p, "bin %d not packed in the pool", m.tile);
        CHECK(m.cap % 64 == 0 && m.cap >= bin_capacity(m.used, slack) && m.cap < kBinCapLimit,
              "bin %d: capacity %d for %lld particles", m.tile, m.cap, (long long)m.used);
        CHECK(!bin_listed(m.cap, m.used, frac), "bin %d still listed after its move", m.tile);
        p += m.cap;
    }
    CHECK(p == pool && pool <= alloc, "pool end %lld, alloc %lld", (long long)pool,
          (long long)alloc);
}

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 20000;
    std::mt19937_64 rng(12345);
    int counts[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) one_case(i, rng, counts);
    std::printf("none %d moves %d whole %d toobig %d failures %d\n", counts[0], counts[1],
                counts[2], counts[3], failures);
    // a bin whose capacity was just set is neither tight nor listed, with the
    // slack (a full re-bin) or twice the slack (a move)
    for (int64_t c = 0; c < 200000; c += 7)
        for (double s : {0.01, 0.05, 0.15, 0.3, 0.6})
            for (double k : {1.0, 2.0})
                if (bin_listed(bin_capacity(c, k * s), c, tight_fraction(s))) {
                    std::printf("fresh bin of %lld at slack %g x%g is listed\n", (long long)c, s,
                                k);
                    return 1;
                }
    return failures ? 1 : 0;
}
