// ohpic_api.cpp -- host runtime and C ABI of libohpic.so (see include/ohpic.h).
//
// A context owns one simulation instance: its device buffers, one HIP stream,
// the tile bins of every species and the work list the particle kernels
// consume.  The per-step sequence is
//   [deposit if the charge grid is stale] -> rho from int64 grid -> 5 FFT
//   passes (x R2HC, y R2HC, fused z forward/divide/inverse, y HC2R, x HC2R)
//   -> node-centred E -> fused gather/push/wrap/deposit/re-bin kernel
// followed by a check of the overflow flags and, when a bin overflowed, a
// full re-bin (repack) of the affected species.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/ohpic.h"
#include "ohpic_fft.hpp"
#include "ohpic_internal.hpp"
#include "ohpic_kernels.hpp"

using namespace ohpic;

namespace {

std::mutex g_err_mu;
std::string g_err;

// A group captures its launch sequence into a hipGraph on its own stream.
// Another thread's HIP calls during that capture can invalidate it (a free or
// a synchronous copy synchronises with every stream of the device) and then
// fail themselves.  Every entry point that makes HIP calls holds this gate
// shared for its duration; a group takes it exclusively for a capture only
// when no other thread holds it (try_lock, so a capture never waits) and
// otherwise launches that sequence directly.  Entry points wait at most for
// one capture (enqueueing the sequence and instantiating its graph, no GPU
// work).
std::shared_mutex g_gate;
thread_local bool t_gate_held = false;

struct GateShared {
    bool owner;
    GateShared() : owner(!t_gate_held)
    {
        if (owner) {
            g_gate.lock_shared();
            t_gate_held = true;
        }
    }
    ~GateShared()
    {
        if (owner) {
            t_gate_held = false;
            g_gate.unlock_shared();
        }
    }
    GateShared(const GateShared &) = delete;
    GateShared &operator=(const GateShared &) = delete;
};

// exclusive hold for a capture, trading the caller's shared hold for it;
// false (the shared hold kept) when another thread is inside an entry point
bool gate_capture_begin()
{
    if (t_gate_held) g_gate.unlock_shared();
    if (g_gate.try_lock()) return true;
    if (t_gate_held) g_gate.lock_shared();
    return false;
}

void gate_capture_end()
{
    g_gate.unlock();
    if (t_gate_held) g_gate.lock_shared();
}

constexpr int kNumTimers = 6;
const char *kTimerNames[kNumTimers] = {"push_fused", "field_solve", "deposit", "repack",
                                       "rebalance", "compact"};

struct TimerSlot {
    hipEvent_t a = nullptr, b = nullptr;
    bool pending = false;
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct SpeciesHost {
    int64_t capacity = 0;  // max particles
    int64_t alloc = 0;     // elements per buffer component
    int64_t n = 0;         // particles loaded
    std::vector<int64_t> base;
    std::vector<int32_t> cap;
    int64_t pool = 0;      // first free element of the relocation pool
};

}  // namespace

struct ohpic_ctx {
    ohpic_config cfg{};
    int device = 0;
    hipStream_t st = nullptr;
    int n[3]{}, T[3]{}, nt[3]{};
    int ntiles = 0;
    int64_t ng = 0;
    int charge_bits = 0;
    SpeciesHost sp[kMaxSpecies];
    InstDev h{};                 // host mirror of the device instance record
    InstDev *d_inst = nullptr;
    WorkItem *d_items = nullptr;
    int nitems = 0;
    int nitems_alloc = 0;
    int64_t *d_tmp64 = nullptr;  // [ntiles] scratch for repack / store
    int32_t *d_tmp32a = nullptr; // [ntiles]
    int32_t *d_tmp32b = nullptr; // [ntiles]
    int64_t *d_tmp64b = nullptr; // [ntiles]
    unsigned long long *d_cursor = nullptr;
    int32_t *d_cellcounts = nullptr;
    int *d_par01 = nullptr;      // {0, 1}: d_par01 + par is this instance's parity
    FftAxisDev ax[3]{};
    double *d_lam[3]{};
    FftJob job{};                // Poisson transform of rho into phi
    FftJob *d_job = nullptr;
    int64_t fft_grp[3]{};        // workgroups per axis pass
    int fft_rs[3]{};             // radix set per axis
    std::vector<WorkItem> items;  // host copy of d_items
    uint64_t inst_ver = 0;       // bumped whenever h is uploaded
    uint64_t items_ver = 0;      // bumped whenever the work list changes
    ohpic_group *group = nullptr;
    bool failed = false;         // a step failed: the state is no longer valid
    unsigned long long *h_flags = nullptr;  // pinned
    double *h_diag = nullptr;               // pinned
    int par = 0;
    bool rho_valid = false;
    bool binned = true;
    int64_t step = 0;
    int64_t repacks = 0;
    int64_t rebalances = 0;
    int64_t compactions = 0;
    // listed bins read from the device list (at most kTightListCap); beyond
    // it the bins are found from the downloaded counts.  OHPIC_TIGHT_LIST_LIMIT
    // lowers it, so that tests reach that path at small sizes.
    unsigned long long list_limit = kTightListCap;
    int64_t fast = 0;
    int num_cus = 256;
    bool timing = false;
    TimerSlot timers[kNumTimers];
    std::vector<std::pair<void *, int64_t>> allocs;  // device pointer, bytes
    int64_t bytes = 0;
    std::string err;
};

namespace {

int fail(ohpic_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, OHPIC_EHIP, "%s failed: %s (%s:%d)", #expr,                   \
                        hipGetErrorString(e_), __FILE__, __LINE__);                         \
    } while (0)

template <typename T>
int dalloc(ohpic_ctx *ctx, T **p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    const size_t bytes = (size_t)count * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, OHPIC_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes,
                    hipGetErrorString(e));
    }
    ctx->allocs.emplace_back((void *)*p, (int64_t)bytes);
    ctx->bytes += (int64_t)bytes;
    return OHPIC_OK;
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

double config_slack(const ohpic_config &c) { return c.slack > 0 ? c.slack : 0.15; }

// Elements per buffer component: the bins of capacity particles with their
// slack and per-bin round-ups, and the relocation pool of tight bins.
int64_t species_alloc(const ohpic_config &c, int s, int ntiles)
{
    const double slack = config_slack(c);
    return (int64_t)std::ceil((double)c.capacity[s] * (1.0 + slack * (1.0 + kPoolSlackShare))) +
           (int64_t)ntiles * 80 + 64;
}


// Default tile: 8^3 cells when the grid holds at least 512 such tiles, else
// 4^3 (fewer, larger tiles leave CUs idle: at 32^3 and 48^3 the 4^3 tiles
// step 1.2-2.1x faster, from 96^3 up the 8^3 tiles 1.1-1.9x).
void pick_tile(const ohpic_config &c, int T[3])
{
    for (int a = 0; a < 3; ++a) T[a] = c.tile[a];
    if (T[0] <= 0 || T[1] <= 0 || T[2] <= 0) {
        const int64_t tiles8 = (int64_t)((c.nx + 7) / 8) * ((c.ny + 7) / 8) * ((c.nz + 7) / 8);
        const int t = tiles8 >= 512 ? 8 : 4;
        T[0] = T[1] = T[2] = t;
    }
}

int derive_charge_bits(const ohpic_config &c, int64_t ng)
{
    if (c.charge_bits > 0) return c.charge_bits;
    // (1) every contribution |q w| * 2^S < 2^51 (exact rounding in fix_round)
    // (2) node sums up to 1024x the mean species density stay below 2^62
    double qmax = 0.0, bound = 0.0;
    for (int s = 0; s < c.nspecies; ++s) {
        qmax = std::max(qmax, std::fabs(c.q_macro[s]));
        bound += std::fabs(c.q_macro[s]) * (double)c.capacity[s] / (double)std::max<int64_t>(ng, 1);
    }
    if (qmax <= 0) qmax = 1.0;
    bound = std::max(bound * 1024.0, 1.0);
    const int s1 = 50 - (int)std::ceil(std::log2(qmax));
    const int s2 = 62 - (int)std::ceil(std::log2(bound));
    return std::max(8, std::min(s1, s2));
}

// Work items a context builds (build_work_items): chunks per tile given, or
// the automatic choice, which never exceeds 8 items per CU plus one per tile.
int64_t max_work_items(const ohpic_config &c, int64_t ntiles, int num_cus)
{
    if (c.chunks > 0) return ntiles * c.chunks;
    return ntiles + (int64_t)num_cus * 8;
}

// Device bytes ohpic_create allocates for c (plus the lazily allocated
// per-cell count buffer), allocation by allocation.
int64_t estimate_bytes(const ohpic_config &c, int num_cus)
{
    int T[3];
    pick_tile(c, T);
    const int64_t nt = (int64_t)((c.nx + T[0] - 1) / T[0]) * ((c.ny + T[1] - 1) / T[1]) *
                       ((c.nz + T[2] - 1) / T[2]);
    const int64_t ng = (int64_t)c.nx * c.ny * c.nz;
    const int64_t per_particle = (c.precision == OHPIC_MIXED ? 7 * 4 : 6 * 8) + (c.track_ids ? 8 : 0);
    int64_t b = 0;
    for (int s = 0; s < c.nspecies; ++s) {
        b += 2 * std::max<int64_t>(species_alloc(c, s, (int)nt), 1) * per_particle;
        b += nt * (2 * 8 + 8 + 4);  // bin words x2, bin base, bin capacity
    }
    b += ng * (2 * 8 + 5 * 8 + 4);           // int64 charge x2, rho, phi, E, cell counts
    b += nt * (8 + 8 + 4 + 4) + 8;           // host-transfer scratch, cursor
    b += 3 * kEfieldMaxBlocks * 8;           // field-energy partials
    b += kDiagSlots * 8 + kFlagWords * 8 + kTightListCap * 8 + (int64_t)sizeof(InstDev) + 2 * 4 +
         (int64_t)sizeof(FftJob);
    b += (int64_t)(c.nx + c.ny + c.nz) * (16 + 8);  // twiddles, Green's-function tables
    b += max_work_items(c, nt, num_cus) * ((int64_t)sizeof(WorkItem) + kKeSlots * 8);
    return b;
}

int sync_upload_inst(ohpic_ctx *ctx)
{
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_inst, &ctx->h, sizeof(InstDev), hipMemcpyHostToDevice,
                               ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ctx->inst_ver += 1;
    return OHPIC_OK;
}

const int *parity_of(const ohpic_ctx *ctx) { return ctx->d_par01 + ctx->par; }

void timer_begin(ohpic_ctx *ctx, int k)
{
    if (!ctx->timing) return;
    TimerSlot &t = ctx->timers[k];
    if (!t.a) {
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
    }
    if (t.pending) {
        (void)hipEventSynchronize(t.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        t.total_ms += ms;
        t.launches += 1;
        t.pending = false;
    }
    (void)hipEventRecord(t.a, ctx->st);
}

void timer_end(ohpic_ctx *ctx, int k)
{
    if (!ctx->timing) return;
    TimerSlot &t = ctx->timers[k];
    (void)hipEventRecord(t.b, ctx->st);
    t.pending = true;
}

void timers_collect(ohpic_ctx *ctx)
{
    for (int k = 0; k < kNumTimers; ++k) {
        TimerSlot &t = ctx->timers[k];
        if (!t.pending) continue;
        (void)hipEventSynchronize(t.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        t.total_ms += ms;
        t.launches += 1;
        t.pending = false;
    }
}

int build_work_items(ohpic_ctx *ctx)
{
    const int nt = ctx->ntiles;
    std::vector<int64_t> load(nt, 0);
    int64_t maxcap = 0;
    for (int s = 0; s < ctx->cfg.nspecies; ++s)
        for (int t = 0; t < nt; ++t) {
            load[t] += ctx->sp[s].cap.empty() ? 0 : ctx->sp[s].cap[t];
            if (!ctx->sp[s].cap.empty()) maxcap = std::max<int64_t>(maxcap, ctx->sp[s].cap[t]);
        }
    int nch = ctx->cfg.chunks;
    if (nch <= 0) {
        const int64_t target = (int64_t)ctx->num_cus * 8;
        nch = (int)std::max<int64_t>(1, (target + nt - 1) / nt);
        // keep at least ~512 particles per chunk of the fullest bin
        const int64_t lim = std::max<int64_t>(1, maxcap / 512);
        nch = (int)std::min<int64_t>(nch, lim);
    }
    std::vector<int> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return load[a] > load[b]; });
    std::vector<WorkItem> items;
    items.reserve((size_t)nt * nch);
    for (int c = 0; c < nch; ++c)
        for (int i = 0; i < nt; ++i)
            items.push_back(WorkItem{0, order[i], (int16_t)c, (int16_t)nch, 0});
    // heavy tiles first within each chunk round: interleave so the first
    // dispatched items are the heaviest chunks
    std::stable_sort(items.begin(), items.end(), [&](const WorkItem &a, const WorkItem &b) {
        return load[a.tile] > load[b.tile];
    });
    for (size_t i = 0; i < items.size(); ++i) items[i].slot = (int32_t)i;
    if ((int)items.size() > ctx->nitems_alloc) {
        for (void *old : {(void *)ctx->d_items, (void *)ctx->h.ke_partial}) {
            if (!old) continue;
            HIPCHK(ctx, hipStreamSynchronize(ctx->st));
            (void)hipFree(old);
            auto it = std::find_if(ctx->allocs.begin(), ctx->allocs.end(),
                                   [&](const std::pair<void *, int64_t> &a) { return a.first == old; });
            ctx->bytes -= it->second;
            ctx->allocs.erase(it);
        }
        ctx->d_items = nullptr;
        ctx->h.ke_partial = nullptr;
        int rc = dalloc(ctx, &ctx->d_items, (int64_t)items.size());
        if (rc) return rc;
        rc = dalloc(ctx, &ctx->h.ke_partial, (int64_t)items.size() * kKeSlots);
        if (rc) return rc;
        HIPCHK(ctx, launch_fill_f64(ctx->h.ke_partial, (int64_t)items.size() * kKeSlots, 0.0,
                                    ctx->st));
        ctx->nitems_alloc = (int)items.size();
    }
    ctx->nitems = (int)items.size();
    ctx->h.nitems = ctx->nitems;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_items, items.data(), items.size() * sizeof(WorkItem),
                               hipMemcpyHostToDevice, ctx->st));
    ctx->items = std::move(items);
    ctx->items_ver += 1;
    return sync_upload_inst(ctx);
}

// Assign capacities/bases from per-bin counts and upload them.
int set_bins(ohpic_ctx *ctx, int s, const std::vector<int64_t> &counts)
{
    SpeciesHost &S = ctx->sp[s];
    const double slack = config_slack(ctx->cfg);
    const int nt = ctx->ntiles;
    S.base.assign(nt, 0);
    S.cap.assign(nt, 0);
    int64_t off = 0;
    for (int t = 0; t < nt; ++t) {
        const int64_t cap = bin_capacity(counts[t], slack);
        if (cap >= kBinCapLimit)
            return fail(ctx, OHPIC_ENOMEM, "bin %d needs %lld slots (limit 2^29): use smaller tiles",
                        t, (long long)cap);
        S.base[t] = off;
        S.cap[t] = (int32_t)cap;
        off += cap;
    }
    if (off > S.alloc)
        return fail(ctx, OHPIC_ENOMEM,
                    "species %d: bins need %lld elements, buffer holds %lld (raise capacity)", s,
                    (long long)off, (long long)S.alloc);
    S.pool = off;
    HIPCHK(ctx, hipMemcpyAsync(ctx->h.sp[s].bin_base, S.base.data(), nt * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h.sp[s].bin_cap, S.cap.data(), nt * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->st));
    return OHPIC_OK;
}

int read_flags(ohpic_ctx *ctx)
{
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_flags, ctx->h.flags, kFlagSlots * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return OHPIC_OK;
}

int clear_flags(ohpic_ctx *ctx)
{
    HIPCHK(ctx, launch_fill_u64(ctx->h.flags, kFlagWords, 0ull, ctx->st));
    return OHPIC_OK;
}

int download_words(ohpic_ctx *ctx, const unsigned long long *dev, std::vector<int64_t> &F,
                   std::vector<int64_t> &B)
{
    const int nt = ctx->ntiles;
    std::vector<unsigned long long> w(nt);
    HIPCHK(ctx, hipMemcpyAsync(w.data(), dev, nt * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    F.resize(nt);
    B.resize(nt);
    for (int t = 0; t < nt; ++t) {
        F[t] = w[t] & 0xffffffffull;
        B[t] = (int64_t)(w[t] >> 32);
    }
    return OHPIC_OK;
}

// Bin a contiguous list of n particles held in buffer par^1 into the bins of
// buffer par (wrap != 0: wrap positions periodically first).
int bin_from_staging(ohpic_ctx *ctx, int s, int64_t n, int wrap)
{
    const int nt = ctx->ntiles;
    const int lbuf = ctx->par ^ 1;
    unsigned long long *cnt = ctx->h.sp[s].count[ctx->par];
    HIPCHK(ctx, launch_fill_u64(cnt, nt, 0ull, ctx->st));
    int rc = clear_flags(ctx);
    if (rc) return rc;
    if (ctx->cfg.precision == OHPIC_MIXED)
        HIPCHK(ctx, launch_bin_count_mx(ctx->d_inst, 0, s, lbuf, n, ctx->T, ctx->st));
    else
        HIPCHK(ctx, launch_bin_count(ctx->d_inst, 0, s, lbuf, n, ctx->T, wrap, ctx->st));
    std::vector<int64_t> F, B;
    if ((rc = download_words(ctx, cnt, F, B))) return rc;
    if ((rc = read_flags(ctx))) return rc;
    if (ctx->h_flags[kFlagRange])
        return fail(ctx, OHPIC_EINVAL, "species %d: %llu particle positions outside [0, n)", s,
                    (unsigned long long)ctx->h_flags[kFlagRange]);
    rc = set_bins(ctx, s, F);
    if (rc) return rc;
    HIPCHK(ctx, launch_fill_u64(cnt, nt, 0ull, ctx->st));
    if (ctx->cfg.precision == OHPIC_MIXED)
        HIPCHK(ctx, launch_bin_scatter_mx(ctx->d_inst, 0, s, lbuf, n, ctx->T, ctx->st));
    else
        HIPCHK(ctx, launch_bin_scatter(ctx->d_inst, 0, s, lbuf, n, ctx->T, wrap, ctx->st));
    rc = read_flags(ctx);
    if (rc) return rc;
    if (ctx->h_flags[kFlagOverflow])
        return fail(ctx, OHPIC_ENOMEM, "binning overflowed (internal error)");
    ctx->sp[s].n = n;
    return OHPIC_OK;
}

// Compact both runs of every bin of buffer par into a contiguous list in
// buffer par^1; returns the number of particles listed.
int bins_to_staging(ohpic_ctx *ctx, int s, int64_t *nlisted)
{
    const int nt = ctx->ntiles;
    std::vector<int64_t> F, B;
    int rc = download_words(ctx, ctx->h.sp[s].count[ctx->par], F, B);
    if (rc) return rc;
    std::vector<int64_t> off(nt);
    int64_t tot = 0;
    for (int t = 0; t < nt; ++t) {
        off[t] = tot;
        const int64_t cap = ctx->sp[s].cap[t];
        const int64_t f = std::min(F[t], cap);
        tot += f + std::min(B[t], cap - f);
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp64, off.data(), nt * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->st));
    if (ctx->cfg.precision == OHPIC_MIXED)
        HIPCHK(ctx, launch_bins_to_list_mx(ctx->d_inst, 0, s, ctx->par, ctx->par ^ 1, nullptr, 0, nt,
                                           ctx->d_tmp64, ctx->st));
    else
        HIPCHK(ctx, launch_bins_to_list(ctx->d_inst, 0, s, ctx->par, ctx->par ^ 1, nullptr, 0, nt,
                                        ctx->d_tmp64, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    *nlisted = tot;
    return OHPIC_OK;
}

int zero_next_counts(ohpic_ctx *ctx)
{
    for (int s = 0; s < ctx->cfg.nspecies; ++s)
        HIPCHK(ctx, launch_fill_u64(ctx->h.sp[s].count[ctx->par ^ 1], ctx->ntiles, 0ull, ctx->st));
    return OHPIC_OK;
}

// Moves the listed bins of species s (cands) into the relocation pool with
// fresh slack (plan_relocation), or only its tight ones when the pool cannot
// take them all: their particles are compacted into the pool slots within
// buffer par (from their old slots, which stay unused until the next full
// re-bin), then the bins take their new bases, capacities and counts.
// Returns 1 when the pool cannot take even the tight bins, so that the caller
// compacts every species into the other buffer (fresh pools).
int relocate_bins(ohpic_ctx *ctx, int s, std::vector<BinCand> cands)
{
    SpeciesHost &S = ctx->sp[s];
    std::vector<BinMove> moves;
    int64_t pool = S.pool;
    int big = -1;
    const double slack = config_slack(ctx->cfg);
    int plan = plan_relocation(cands, slack, &pool, S.alloc, moves, &big);
    if (plan == kPlanWhole) {
        cands.erase(std::remove_if(cands.begin(), cands.end(),
                                   [&](const BinCand &c) {
                                       return !bin_tight(S.cap[c.tile], c.used,
                                                         ctx->h.tight_frac);
                                   }),
                    cands.end());
        if (!cands.empty()) plan = plan_relocation(cands, slack, &pool, S.alloc, moves, &big);
        if (cands.empty()) return OHPIC_OK;
    }
    if (plan == kPlanWhole) return 1;
    if (plan == kPlanTooBig)
        return fail(ctx, OHPIC_ENOMEM, "bin %d needs 2^29 slots or more: use smaller tiles", big);
    const int nm = (int)moves.size();
    std::vector<int32_t> tiles(nm), cap(nm);
    std::vector<int64_t> base(nm), used(nm);
    for (int k = 0; k < nm; ++k) {
        tiles[k] = moves[k].tile;
        base[k] = moves[k].base;
        cap[k] = moves[k].cap;
        used[k] = moves[k].used;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp32a, tiles.data(), nm * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp64, base.data(), nm * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp32b, cap.data(), nm * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp64b, used.data(), nm * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->st));
    const int par = ctx->par;
    // the particles first (the compaction reads the old bases), then the layout
    if (ctx->cfg.precision == OHPIC_MIXED)
        HIPCHK(ctx, launch_bins_to_list_mx(ctx->d_inst, 0, s, par, par, ctx->d_tmp32a, 0, nm,
                                           ctx->d_tmp64, ctx->st));
    else
        HIPCHK(ctx, launch_bins_to_list(ctx->d_inst, 0, s, par, par, ctx->d_tmp32a, 0, nm,
                                        ctx->d_tmp64, ctx->st));
    HIPCHK(ctx, launch_set_bins(ctx->d_inst, 0, s, par, ctx->d_tmp32a, ctx->d_tmp64,
                                ctx->d_tmp32b, ctx->d_tmp64b, nm, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    for (const BinMove &m : moves) {
        S.base[m.tile] = m.base;
        S.cap[m.tile] = m.cap;
    }
    S.pool = pool;
    return OHPIC_OK;
}

// Compacts every species into buffer par^1, each bin laid out as a full
// re-bin lays it out for its count (one read and one write of every particle;
// no particle changes bin), then makes that buffer the current one: the
// charge grids swap with it, and every relocation pool starts empty.
int compact_bins(ohpic_ctx *ctx)
{
    timer_begin(ctx, 5);
    const int nt = ctx->ntiles, par = ctx->par;
    const double slack = config_slack(ctx->cfg);
    int rc;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        if ((int)ctx->sp[s].cap.size() != nt) continue;
        std::vector<int64_t> F, B;
        if ((rc = download_words(ctx, ctx->h.sp[s].count[par], F, B))) return rc;
        std::vector<int64_t> used(nt), off(nt);
        std::vector<unsigned long long> words(nt);
        int64_t o = 0;
        for (int t = 0; t < nt; ++t) {
            used[t] = F[t] + B[t];
            words[t] = (unsigned long long)used[t];  // one front run
            off[t] = o;
            o += bin_capacity(used[t], slack);
        }
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp64, off.data(), nt * sizeof(int64_t),
                                   hipMemcpyHostToDevice, ctx->st));
        // the copy reads the current layout; set_bins then uploads the new one
        // (the same bases as off)
        if (ctx->cfg.precision == OHPIC_MIXED)
            HIPCHK(ctx, launch_bins_to_list_mx(ctx->d_inst, 0, s, par, par ^ 1, nullptr, 0, nt,
                                               ctx->d_tmp64, ctx->st));
        else
            HIPCHK(ctx, launch_bins_to_list(ctx->d_inst, 0, s, par, par ^ 1, nullptr, 0, nt,
                                            ctx->d_tmp64, ctx->st));
        if ((rc = set_bins(ctx, s, used))) return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->h.sp[s].count[par ^ 1], words.data(),
                                   nt * sizeof(unsigned long long), hipMemcpyHostToDevice,
                                   ctx->st));
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    }
    ctx->par ^= 1;
    std::swap(ctx->h.rho_acc[0], ctx->h.rho_acc[1]);
    timer_end(ctx, 5);
    ctx->compactions += 1;
    if ((rc = sync_upload_inst(ctx))) return rc;
    return build_work_items(ctx);
}

// After a step that left a bin tight: move the bins k_end_step_b listed
// (`listed` of them, from tight_list; when more than list_limit, the listed
// bins are found from every species' counts).  When a species' pool
// cannot take its tight bins, every species is compacted instead.
int rebalance_bins(ohpic_ctx *ctx, unsigned long long listed)
{
    timer_begin(ctx, 4);
    const int ns = ctx->cfg.nspecies, nt = ctx->ntiles;
    std::vector<std::vector<BinCand>> cands(ns);
    int rc;
    if (listed <= ctx->list_limit) {
        std::vector<unsigned long long> e(listed);
        HIPCHK(ctx, hipMemcpyAsync(e.data(), ctx->h.tight_list, listed * sizeof(e[0]),
                                   hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
        for (unsigned long long v : e)
            cands[tight_entry_species(v)].push_back({tight_entry_tile(v), tight_entry_used(v)});
        // tile order: the pool layout does not depend on the scan's order
        for (auto &c : cands)
            std::sort(c.begin(), c.end(),
                      [](const BinCand &a, const BinCand &b) { return a.tile < b.tile; });
    } else {
        for (int s = 0; s < ns; ++s) {
            if ((int)ctx->sp[s].cap.size() != nt) continue;
            std::vector<int64_t> F, B;
            if ((rc = download_words(ctx, ctx->h.sp[s].count[ctx->par], F, B))) return rc;
            for (int t = 0; t < nt; ++t)
                if (bin_listed(ctx->sp[s].cap[t], F[t] + B[t], ctx->h.tight_frac))
                    cands[s].push_back({t, F[t] + B[t]});
        }
    }
    bool full = false;
    for (int s = 0; s < ns && !full; ++s) {
        if (cands[s].empty()) continue;
        rc = relocate_bins(ctx, s, cands[s]);
        if (rc < 0) return rc;
        full = rc == 1;
    }
    timer_end(ctx, 4);
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ctx->rebalances += 1;
    // a relocation pool ran out: compact every species (fresh pools)
    return full ? compact_bins(ctx) : OHPIC_OK;
}

// Retry of a step whose exchange overflowed a bin: run the kernel again in
// list mode (every particle to a contiguous list of buffer par^1), then bin
// each species' list into buffer par with capacities sized from it.
int rerun_as_list(ohpic_ctx *ctx, int list_mode)
{
    timer_begin(ctx, 3);
    int rc;
    HIPCHK(ctx, launch_prepare_next_b(ctx->d_inst, parity_of(ctx), 1, ctx->ng, ctx->ntiles,
                                      nullptr, ctx->st));
    HIPCHK(ctx, launch_particles(list_mode, ctx->cfg.precision, ctx->cfg.weight_order, ctx->T,
                                 ctx->d_inst, ctx->d_items, ctx->nitems, parity_of(ctx), ctx->st));
    if (list_mode == kModeFusedList) HIPCHK(ctx, launch_reduce_ke_b(ctx->d_inst, 1, ctx->st));
    if ((rc = read_flags(ctx))) return rc;
    if (ctx->h_flags[kFlagRange])
        return fail(ctx, OHPIC_ESTATE,
                    "%llu particles moved too far (more than two box lengths; in mixed "
                    "precision 64 cells) or have non-finite coordinates",
                    (unsigned long long)ctx->h_flags[kFlagRange]);
    std::vector<int64_t> nlist(ctx->cfg.nspecies);
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        unsigned long long w = 0;
        HIPCHK(ctx, hipMemcpyAsync(&w, ctx->h.sp[s].count[ctx->par ^ 1], sizeof w,
                                   hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
        nlist[s] = (int64_t)w;
    }
    for (int s = 0; s < ctx->cfg.nspecies; ++s)
        if ((rc = bin_from_staging(ctx, s, nlist[s], 0))) return rc;
    if (list_mode == kModeFusedList) {
        // the step's charge is in accumulator par^1 while the particles are
        // back in buffer par: swap the accumulators
        std::swap(ctx->h.rho_acc[0], ctx->h.rho_acc[1]);
        if ((rc = sync_upload_inst(ctx))) return rc;
    }
    timer_end(ctx, 3);
    if ((rc = clear_flags(ctx))) return rc;
    ctx->repacks += 1;
    return build_work_items(ctx);
}

int do_deposit(ohpic_ctx *ctx)
{
    timer_begin(ctx, 2);
    HIPCHK(ctx, launch_fill_u64(ctx->h.rho_acc[ctx->par], ctx->ng, 0ull, ctx->st));
    HIPCHK(ctx, launch_particles(kModeDeposit, ctx->cfg.precision, ctx->cfg.weight_order, ctx->T,
                                 ctx->d_inst, ctx->d_items, ctx->nitems, parity_of(ctx), ctx->st));
    timer_end(ctx, 2);
    ctx->rho_valid = true;
    return OHPIC_OK;
}

// Poisson solve of the instance table `insts` (n instances, FFT jobs `jobs`)
// from rho: five axis passes, node E and its energy partials (reduced when
// reduce_e).
hipError_t solve_passes(const FftJob *jobs, const InstDev *insts, int n, const int64_t grp[3],
                        const int rs[3], int z256, int max_eblocks, bool reduce_e,
                        hipStream_t st)
{
    hipError_t e = fft_pass(kPassR2hcXFromSrc, jobs, n, grp[0], rs[0], st);
    if (e == hipSuccess) e = fft_pass(kPassR2hcY, jobs, n, grp[1], rs[1], st);
    if (e == hipSuccess) e = fft_poisson_z(jobs, n, grp[2], rs[2], z256, st);
    if (e == hipSuccess) e = fft_pass(kPassHc2rY, jobs, n, grp[1], rs[1], st);
    if (e == hipSuccess) e = fft_pass(kPassHc2rXScaled, jobs, n, grp[0], rs[0], st);
    if (e == hipSuccess) e = launch_efield_b(insts, n, max_eblocks, reduce_e, st);
    return e;
}

// The field solve from rho (FFT passes, E); reduce_e: also the field-energy
// sums (a step leaves them to its end-of-step reduction).
int do_solve(ohpic_ctx *ctx, bool reduce_e)
{
    timer_begin(ctx, 1);
    HIPCHK(ctx, solve_passes(ctx->d_job, ctx->d_inst, 1, ctx->fft_grp, ctx->fft_rs,
                             ctx->n[2] == 256 ? kZ256All : kZ256None, efield_blocks_of(ctx->n[0], ctx->n[1], ctx->n[2]), reduce_e,
                             ctx->st));
    timer_end(ctx, 1);
    return OHPIC_OK;
}

// Flags of a finished re-binning step: counts fast particles, fails on
// out-of-range movers; 1 when a bin overflowed.
int absorb_flags(ohpic_ctx *ctx, const unsigned long long *f)
{
    if (f[kFlagRange])
        return fail(ctx, OHPIC_ESTATE,
                    "%llu particles moved too far in one step (more than two box lengths; in "
                    "mixed precision 64 cells) or have non-finite coordinates",
                    (unsigned long long)f[kFlagRange]);
    ctx->fast += (int64_t)f[kFlagFast];
    return f[kFlagOverflow] ? 1 : 0;
}

int one_step(ohpic_ctx *ctx)
{
    int rc;
    if (!ctx->rho_valid && (rc = do_deposit(ctx))) return rc;
    // rho from the charge grid and the next buffer's zeroing in one sweep,
    // the solve from rho, the step, and all energy sums in one reduction
    HIPCHK(ctx, launch_begin_step_b(ctx->d_inst, parity_of(ctx), 1, ctx->ng, ctx->ntiles,
                                    nullptr, ctx->st));
    if ((rc = do_solve(ctx, false))) return rc;
    timer_begin(ctx, 0);
    HIPCHK(ctx, launch_particles(kModeFused, ctx->cfg.precision, ctx->cfg.weight_order, ctx->T,
                                 ctx->d_inst, ctx->d_items, ctx->nitems, parity_of(ctx), ctx->st));
    timer_end(ctx, 0);
    // the parity word is only read here (no out: no flip)
    HIPCHK(ctx, launch_end_step_b(ctx->d_inst, ctx->d_par01 + ctx->par, 1, nullptr, ctx->st));
    if ((rc = read_flags(ctx))) return rc;
    timers_collect(ctx);
    rc = absorb_flags(ctx, ctx->h_flags);
    if (rc < 0) return rc;
    if (rc == 1) {
        // a bin overflowed: redo the step into a list and bin it into buffer par
        if ((rc = rerun_as_list(ctx, kModeFusedList))) return rc;
    } else {
        ctx->par ^= 1;
    }
    ctx->step += 1;
    ctx->rho_valid = true;
    // the step left a bin short of free slots: restore its slack now, before
    // an exchange overflows it
    if (rc == 0 && ctx->h_flags[kFlagTight])
        return rebalance_bins(ctx, ctx->h_flags[kFlagListed]);
    return OHPIC_OK;
}

int grid_copy_out(ohpic_ctx *ctx, const double *dev, std::vector<double> &host)
{
    host.resize(ctx->ng);
    HIPCHK(ctx, hipMemcpyAsync(host.data(), dev, ctx->ng * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return OHPIC_OK;
}

int make_axis(ohpic_ctx *ctx, int n, FftAxisDev &ax)
{
    ax.n = n;
    if (!fft_factorize(n, ax.fac, &ax.nfac))
        return fail(ctx, OHPIC_EINVAL, "axis length %d: unsupported (max %d, prime factors <= %d)",
                    n, kMaxFftLen, kMaxRadix);
    fft_plan_divisors(ax);
    std::vector<double2> tw(n);
    fft_twiddles(n, tw.data());
    double2 *d = nullptr;
    int rc = dalloc(ctx, &d, n);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(d, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ax.tw = d;
    return OHPIC_OK;
}

int group_detach(ohpic_group *g, ohpic_ctx *ctx);

void release(ohpic_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->group) (void)group_detach(ctx->group, ctx);
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    for (auto &a : ctx->allocs) (void)hipFree(a.first);
    ctx->allocs.clear();
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    if (ctx->h_diag) (void)hipHostFree(ctx->h_diag);
    for (auto &t : ctx->timers) {
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
    }
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

}  // namespace

// FP64 positions/velocities <-> the mixed-precision list form: per particle
// the tile index, per axis a word holding the cell relative to the tile
// origin (top 8 bits) and the in-cell offset as a 24-bit fraction, and FP32
// velocities.  The cell is the FP64 int() truncation and an offset that
// rounds up to 1 is clamped below it, so a particle never changes cell in the
// conversion.  Positions outside [0, n) get tile -1, which the binning kernel
// rejects.
void to_mixed(ohpic_ctx *ctx, int64_t n, const double *const src[6], std::vector<int32_t> &tile,
              std::vector<uint32_t> word[3], std::vector<float> vel[3])
{
    tile.resize(n);
    for (int a = 0; a < 3; ++a) {
        word[a].resize(n);
        vel[a].resize(n);
    }
    for (int64_t i = 0; i < n; ++i) {
        int tc[3];
        bool bad = false;
        for (int a = 0; a < 3; ++a) {
            const double p = src[a][i];
            const int na = ctx->n[a];
            if (!(p >= 0.0 && p < (double)na)) {
                bad = true;
                tc[a] = 0;
                word[a][i] = 0u;
                continue;
            }
            const int c = (int)p;
            const double f = std::min(std::rint((p - (double)c) * 16777216.0), 16777215.0);
            tc[a] = c / ctx->T[a];
            word[a][i] = ((uint32_t)(c - tc[a] * ctx->T[a]) << 24) | (uint32_t)f;
        }
        tile[i] = bad ? -1 : tc[0] + ctx->nt[0] * (tc[1] + ctx->nt[1] * tc[2]);
        for (int a = 0; a < 3; ++a) vel[a][i] = (float)src[3 + a][i];
    }
}

void from_mixed(const ohpic_ctx *ctx, int64_t n, const std::vector<int32_t> &tile,
                const std::vector<uint32_t> word[3], const std::vector<float> vel[3],
                double *const dst[6])
{
    for (int64_t i = 0; i < n; ++i) {
        const int t = tile[i];
        const int tc[3] = {t % ctx->nt[0], (t / ctx->nt[0]) % ctx->nt[1],
                           t / (ctx->nt[0] * ctx->nt[1])};
        for (int a = 0; a < 3; ++a) {
            const uint32_t w = word[a][i];
            const int c = tc[a] * ctx->T[a] + ((int32_t)w >> 24);
            if (dst[a]) dst[a][i] = (double)c + (double)(w & 0xffffffu) * 0x1p-24;
            if (dst[3 + a]) dst[3 + a][i] = (double)vel[a][i];
        }
    }
}

// ============================================================== FFT helpers
namespace ohpic {

void set_global_error(const char *msg)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg ? msg : "";
}

bool fft_factorize(int n, int *fac, int *nfac)
{
    if (n < 1 || n > kMaxFftLen) return false;
    int k = 0;
    int m = n;
    while (m % 4 == 0 && k < kMaxFactors) {
        fac[k++] = 4;
        m /= 4;
    }
    for (int p = 2; m > 1 && k < kMaxFactors; ++p) {
        while (m % p == 0 && k < kMaxFactors) {
            if (p > kMaxRadix) return false;
            fac[k++] = p;
            m /= p;
        }
    }
    if (m != 1) return false;
    *nfac = k;
    return true;
}

void fft_twiddles(int n, double2 *tw)
{
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        // reduce to the first octant for accuracy
        const long double a = two_pi * (long double)k / (long double)n;
        tw[k].x = (double)cosl(a);
        tw[k].y = (double)-sinl(a);
    }
}

}  // namespace ohpic

// ================================================================== groups
// Steps per launch sequence of a group: each sequence is one graph launch and
// one host synchronisation (with one read-back of every step's flags).
constexpr int kGroupMaxBatch = 16;
constexpr int kGroupDefaultBatch = 8;

struct ohpic_group {
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<ohpic_ctx *> members;
    std::vector<uint64_t> seen_inst, seen_items;  // member versions in the tables
    std::vector<int> dev_par;                     // parity the device table holds
    InstDev *d_insts = nullptr;
    FftJob *d_jobs = nullptr;
    int *d_parity = nullptr;
    unsigned long long *d_flags = nullptr;
    unsigned long long *h_flags = nullptr;  // pinned
    int cap_insts = 0;
    WorkItem *d_items = nullptr;
    int64_t cap_items = 0;
    int nitems = 0;
    bool dirty = true;
    int64_t max_ng = 0;
    int max_ntiles = 0, max_rows = 0;  // max_rows: largest E-kernel block count
    int z256 = kZ256None;              // members whose Poisson z pass has 256-point lines
    int64_t grp[3]{};
    int rs[3]{};
    int T[3]{};
    int precision = -1, weight_order = -1;
    bool use_graph = true;
    bool persistent = true;          // queued particle launch over a resident grid
    unsigned int *d_queue = nullptr;  // its work counter
    int queue_blocks = 0;             // resident workgroups of that launch
    int batch = kGroupDefaultBatch;  // steps per launch sequence (one host sync each)
    hipGraph_t graph = nullptr;       // one step
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_b = nullptr;     // `batch` steps
    hipGraphExec_t exec_b = nullptr;
    int64_t steps = 0, captures = 0;
    int64_t capture_failures = 0;  // sequences launched directly after a failed capture
    int64_t replayed = 0;          // launch sequences replayed from a graph
    int64_t direct = 0;            // launch sequences enqueued kernel by kernel
    std::string err;
};

namespace {

int gfail(ohpic_group *g, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (g) g->err = buf;
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = buf;
    return code;
}

#define GHIP(g, expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return gfail(g, OHPIC_EHIP, "%s failed: %s (%s:%d)", #expr,                    \
                         hipGetErrorString(e_), __FILE__, __LINE__);                        \
    } while (0)

void drop_graph(ohpic_group *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->exec_b) (void)hipGraphExecDestroy(g->exec_b);
    if (g->graph_b) (void)hipGraphDestroy(g->graph_b);
    g->exec = g->exec_b = nullptr;
    g->graph = g->graph_b = nullptr;
}

template <typename T>
void gfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int group_reserve(ohpic_group *g, int ninst, int64_t nitems)
{
    if (ninst > g->cap_insts) {
        const int cap = std::max(ninst, std::max(16, 2 * g->cap_insts));
        GHIP(g, hipStreamSynchronize(g->st));
        drop_graph(g);
        gfree(g->d_insts);
        gfree(g->d_jobs);
        gfree(g->d_parity);
        gfree(g->d_flags);
        if (g->h_flags) (void)hipHostFree(g->h_flags);
        g->h_flags = nullptr;
        g->cap_insts = 0;
        GHIP(g, hipMalloc((void **)&g->d_insts, cap * sizeof(InstDev)));
        GHIP(g, hipMalloc((void **)&g->d_jobs, cap * sizeof(FftJob)));
        GHIP(g, hipMalloc((void **)&g->d_parity, cap * sizeof(int)));
        GHIP(g, hipMalloc((void **)&g->d_flags, (size_t)cap * kGroupMaxBatch * kFlagSlots * 8));
        GHIP(g, hipHostMalloc((void **)&g->h_flags, (size_t)cap * kGroupMaxBatch * kFlagSlots * 8,
                              hipHostMallocDefault));
        g->cap_insts = cap;
    }
    if (nitems > g->cap_items) {
        const int64_t cap = std::max<int64_t>(nitems, std::max<int64_t>(1024, 2 * g->cap_items));
        GHIP(g, hipStreamSynchronize(g->st));
        drop_graph(g);
        gfree(g->d_items);
        g->cap_items = 0;
        GHIP(g, hipMalloc((void **)&g->d_items, cap * sizeof(WorkItem)));
        g->cap_items = cap;
    }
    return OHPIC_OK;
}

// Rebuild the device tables when membership or a member's record / work list
// changed, and upload the parity table when it differs from the members'.
int group_refresh(ohpic_group *g)
{
    const int n = (int)g->members.size();
    bool tables = g->dirty;
    for (int i = 0; i < n && !tables; ++i)
        tables = g->members[i]->inst_ver != g->seen_inst[i] ||
                 g->members[i]->items_ver != g->seen_items[i];
    if (tables) {
        int64_t total = 0;
        for (ohpic_ctx *m : g->members) total += m->nitems;
        if (total > INT32_MAX) return gfail(g, OHPIC_EINVAL, "more than 2^31 work items");
        int rc = group_reserve(g, n, total);
        if (rc) return rc;
        std::vector<InstDev> insts(n);
        std::vector<FftJob> jobs(n);
        std::vector<WorkItem> items;
        items.reserve((size_t)total);
        g->max_ng = 0;
        g->max_ntiles = g->max_rows = 0;
        int n256 = 0;
        for (int a = 0; a < 3; ++a) {
            g->grp[a] = 0;
            g->rs[a] = 0;
        }
        for (int i = 0; i < n; ++i) {
            ohpic_ctx *m = g->members[i];
            insts[i] = m->h;
            jobs[i] = m->job;
            for (WorkItem w : m->items) {
                w.inst = i;
                items.push_back(w);
            }
            g->max_ng = std::max(g->max_ng, m->ng);
            g->max_ntiles = std::max(g->max_ntiles, m->ntiles);
            g->max_rows = std::max(g->max_rows, efield_blocks_of(m->n[0], m->n[1], m->n[2]));
            n256 += m->n[2] == 256;
            for (int a = 0; a < 3; ++a) {
                g->grp[a] = std::max(g->grp[a], m->fft_grp[a]);
                g->rs[a] = std::max(g->rs[a], m->fft_rs[a]);
            }
            g->seen_inst[i] = m->inst_ver;
            g->seen_items[i] = m->items_ver;
        }
        g->z256 = n256 == 0 ? kZ256None : (n256 == n ? kZ256All : kZ256Some);
        // heavy items first over the whole table: the dispatcher then fills
        // the machine with the largest chunks of every instance
        std::stable_sort(items.begin(), items.end(), [&](const WorkItem &a, const WorkItem &b) {
            auto load = [&](const WorkItem &w) {
                const ohpic_ctx *m = g->members[w.inst];
                int64_t l = 0;
                for (int s = 0; s < m->cfg.nspecies; ++s)
                    if (!m->sp[s].cap.empty()) l += m->sp[s].cap[w.tile];
                return l / w.nchunks;
            };
            return load(a) > load(b);
        });
        g->nitems = (int)items.size();
        if (n > 0 && g->queue_blocks == 0)
            g->queue_blocks = particles_queue_blocks(g->precision, g->weight_order, g->T, g->device);
        if (n > 0) {
            GHIP(g, hipMemcpyAsync(g->d_insts, insts.data(), n * sizeof(InstDev),
                                   hipMemcpyHostToDevice, g->st));
            GHIP(g, hipMemcpyAsync(g->d_jobs, jobs.data(), n * sizeof(FftJob),
                                   hipMemcpyHostToDevice, g->st));
        }
        if (!items.empty())
            GHIP(g, hipMemcpyAsync(g->d_items, items.data(), items.size() * sizeof(WorkItem),
                                   hipMemcpyHostToDevice, g->st));
        GHIP(g, hipStreamSynchronize(g->st));
        drop_graph(g);
        g->dirty = false;
    }
    bool par = false;
    for (int i = 0; i < n; ++i) par = par || g->dev_par[i] != g->members[i]->par;
    if (par) {
        std::vector<int> p(n);
        for (int i = 0; i < n; ++i) p[i] = g->members[i]->par;
        GHIP(g, hipMemcpyAsync(g->d_parity, p.data(), n * sizeof(int), hipMemcpyHostToDevice,
                               g->st));
        GHIP(g, hipStreamSynchronize(g->st));
        g->dev_par = p;
    }
    return OHPIC_OK;
}

// One step of every member: solve, fused particle kernel, KE sums, flags.
// Enqueues k consecutive steps of every member; step s leaves the members'
// flags in d_flags[s][member], copied to h_flags in one transfer at the end.
// A member whose bins overflow keeps its parity on the device, so its later
// steps of the sequence repeat that step from the same state.
hipError_t group_enqueue(ohpic_group *g, int k)
{
    const int n = (int)g->members.size();
    hipError_t e = hipSuccess;
    for (int s = 0; s < k && e == hipSuccess; ++s) {
        e = launch_begin_step_b(g->d_insts, g->d_parity, n, g->max_ng, g->max_ntiles,
                                g->persistent ? g->d_queue : nullptr, g->st);
        if (e == hipSuccess)
            e = solve_passes(g->d_jobs, g->d_insts, n, g->grp, g->rs, g->z256, g->max_rows, false,
                             g->st);
        if (e == hipSuccess) {
            if (g->persistent)
                e = launch_particles_queue(g->precision, g->weight_order, g->T, g->d_insts,
                                           g->d_items, g->nitems, g->d_parity, g->d_queue,
                                           g->queue_blocks, g->st);
            else
                e = launch_particles(kModeFused, g->precision, g->weight_order, g->T, g->d_insts,
                                     g->d_items, g->nitems, g->d_parity, g->st);
        }
        if (e == hipSuccess)
            e = launch_end_step_b(g->d_insts, g->d_parity, n,
                                  g->d_flags + (size_t)s * n * kFlagSlots, g->st);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->h_flags, g->d_flags, (size_t)k * n * kFlagSlots * 8,
                           hipMemcpyDeviceToHost, g->st);
    return e;
}

// Captures the k-step sequence (k = 1 or g->batch) into its graph.
int group_capture(ohpic_group *g, int k)
{
    hipGraph_t &graph = k == 1 ? g->graph : g->graph_b;
    hipGraphExec_t &exec = k == 1 ? g->exec : g->exec_b;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    GHIP(g, hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
    const hipError_t e = group_enqueue(g, k);
    hipGraph_t captured = nullptr;
    const hipError_t e2 = hipStreamEndCapture(g->st, &captured);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (captured) (void)hipGraphDestroy(captured);
        return gfail(g, OHPIC_EHIP, "graph capture failed: %s",
                     hipGetErrorString(e != hipSuccess ? e : e2));
    }
    graph = captured;
    GHIP(g, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    g->captures += 1;
    return OHPIC_OK;
}

int group_step(ohpic_group *g, int nsteps)
{
    GHIP(g, hipSetDevice(g->device));
    const int n = (int)g->members.size();
    for (int done = 0; done < nsteps && n > 0;) {
        for (int i = 0; i < n; ++i) {
            ohpic_ctx *m = g->members[i];
            if (!m->binned)
                return gfail(g, OHPIC_ESTATE, "member %d: particles pushed but not re-binned", i);
            if (!m->rho_valid) {
                int rc = do_deposit(m);
                if (rc == OHPIC_OK && hipStreamSynchronize(m->st) != hipSuccess)
                    rc = fail(m, OHPIC_EHIP, "deposit failed");
                if (rc) return gfail(g, rc, "member %d: %s", i, m->err.c_str());
            }
        }
        int rc = group_refresh(g);
        if (rc) return rc;
        // whole sequences of `batch` steps while that many remain, then single
        // steps: one host synchronisation per sequence
        const int k = (g->batch > 1 && nsteps - done >= g->batch) ? g->batch : 1;
        bool have = g->use_graph && (k == 1 ? g->exec : g->exec_b);
        if (g->use_graph && !have && gate_capture_begin()) {
            // no other thread is inside the library: capture now.  A capture
            // that failed enqueued nothing; this sequence then launches its
            // kernels directly and the next one captures again
            have = group_capture(g, k) == OHPIC_OK;
            gate_capture_end();
            if (!have) {
                (void)hipGetLastError();
                g->capture_failures += 1;
            }
        }
        if (have) {
            GHIP(g, hipGraphLaunch(k == 1 ? g->exec : g->exec_b, g->st));
            g->replayed += 1;
        } else {
            // no graph (disabled, or another thread was making HIP calls)
            GHIP(g, group_enqueue(g, k));
            g->direct += 1;
        }
        GHIP(g, hipStreamSynchronize(g->st));
        // every member's steps are accounted for before a failure is reported:
        // failed members leave the group, the others have completed the k steps
        int first_rc = OHPIC_OK;
        std::string msg;
        std::vector<ohpic_ctx *> failed;
        for (int i = 0; i < n; ++i) {
            ohpic_ctx *m = g->members[i];
            // a step whose bins overflowed left the member's state unchanged
            // (the device kept its parity): count the steps that went through,
            // redo the overflowed one as a list and take the rest alone
            int ok = 0;
            bool overflowed = false, last_ok = false;
            rc = OHPIC_OK;
            for (int s = 0; s < k; ++s) {
                const unsigned long long *f = g->h_flags + ((size_t)s * n + i) * kFlagSlots;
                last_ok = false;
                if (f[kFlagOverflow] && !f[kFlagRange]) {
                    if (!overflowed) m->fast += (int64_t)f[kFlagFast];
                    overflowed = true;
                    continue;
                }
                if ((rc = absorb_flags(m, f)) < 0) break;
                ++ok;
                last_ok = true;
            }
            if (rc == OHPIC_OK) {
                if (ok & 1) {
                    m->par ^= 1;
                    g->dev_par[i] ^= 1;
                }
                m->step += ok;
                m->rho_valid = true;
                int r = ok;
                if (r < k && !last_ok) {
                    // the last step of the sequence overflowed: E still belongs
                    // to the member's current state, so that step is redone as
                    // a list (as a single step would be)
                    rc = rerun_as_list(m, kModeFusedList);
                    if (rc == OHPIC_OK) {
                        m->step += 1;
                        ++r;
                    }
                }
                // an overflow followed by a step that went through (bin
                // arrivals race with the front reservation near a full bin)
                // leaves E one step behind: the rest run alone, each with its
                // own field solve
                for (; r < k && rc == OHPIC_OK; ++r) rc = one_step(m);
                // every step of the sequence went through and one left a bin
                // tight: move the bins the last step listed (a member that
                // finished alone or was re-binned has already seen to its bins)
                bool tight = false;
                for (int s = 0; s < k; ++s)
                    tight = tight || g->h_flags[((size_t)s * n + i) * kFlagSlots + kFlagTight];
                if (rc == OHPIC_OK && tight && ok == k)
                    rc = rebalance_bins(
                        m, g->h_flags[((size_t)(k - 1) * n + i) * kFlagSlots + kFlagListed]);
            }
            if (rc < 0) {
                m->failed = true;
                failed.push_back(m);
                if (first_rc == OHPIC_OK) first_rc = rc;
                char buf[64];
                snprintf(buf, sizeof buf, "%smember %d: ", msg.empty() ? "" : "; ", i);
                msg += buf + m->err;
            }
        }
        g->steps += k;
        done += k;
        if (first_rc != OHPIC_OK) {
            for (ohpic_ctx *m : failed) (void)group_detach(g, m);
            return gfail(g, first_rc, "%s (failed members were detached)", msg.c_str());
        }
    }
    return OHPIC_OK;
}

int group_detach(ohpic_group *g, ohpic_ctx *ctx)
{
    auto it = std::find(g->members.begin(), g->members.end(), ctx);
    if (it == g->members.end()) return gfail(g, OHPIC_EINVAL, "context is not a member");
    (void)hipSetDevice(g->device);
    GHIP(g, hipStreamSynchronize(g->st));
    const size_t i = (size_t)(it - g->members.begin());
    g->members.erase(it);
    g->seen_inst.erase(g->seen_inst.begin() + i);
    g->seen_items.erase(g->seen_items.begin() + i);
    g->dev_par.erase(g->dev_par.begin() + i);
    ctx->group = nullptr;
    g->dirty = true;
    // the parity table is indexed by position: upload it again
    for (int &p : g->dev_par) p = -1;
    drop_graph(g);
    return OHPIC_OK;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

int ohpic_config_default(ohpic_config *cfg)
{
    if (!cfg) return fail(nullptr, OHPIC_EINVAL, "cfg is NULL");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->nx = cfg->ny = cfg->nz = 32;
    cfg->nspecies = 2;
    cfg->qm[0] = -1.0;
    cfg->qm[1] = 0.01;
    cfg->q_macro[0] = -1.0 / 16;
    cfg->q_macro[1] = 1.0 / 16;
    cfg->capacity[0] = cfg->capacity[1] = 16LL * 32 * 32 * 32;
    cfg->dt = 0.1;
    cfg->b0[2] = 0.5;
    cfg->precision = OHPIC_FP64;
    cfg->weight_order = OHPIC_WEIGHTS_REFERENCE;
    cfg->slack = 0.15;
    return OHPIC_OK;
}

const char *ohpic_last_error(const ohpic_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    static thread_local std::string copy;
    copy = g_err;
    return copy.c_str();
}

int ohpic_estimate_footprint(const ohpic_config *cfg, int64_t *bytes)
{
    GateShared gate;
    if (!cfg || !bytes) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    if (cfg->chunks < 0 || cfg->chunks > kMaxChunks)
        return fail(nullptr, OHPIC_EINVAL, "chunks must be in [0, %d] (0 = automatic)", kMaxChunks);
    int num_cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess)
        num_cus = prop.multiProcessorCount;
    else
        (void)hipGetLastError();
    *bytes = estimate_bytes(*cfg, num_cus);
    return OHPIC_OK;
}

int ohpic_create(const ohpic_config *cfg, ohpic_ctx **out)
{
    GateShared gate;
    if (!cfg || !out) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    *out = nullptr;
    const ohpic_config &c = *cfg;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1)
        return fail(nullptr, OHPIC_EINVAL, "grid sizes must be >= 1");
    if (c.nspecies < 1 || c.nspecies > OHPIC_MAX_SPECIES)
        return fail(nullptr, OHPIC_EINVAL, "nspecies must be in [1, %d]", OHPIC_MAX_SPECIES);
    if (c.precision != OHPIC_FP64 && c.precision != OHPIC_MIXED)
        return fail(nullptr, OHPIC_EINVAL, "unknown precision mode %d", c.precision);
    if (c.nx > kMaxFftLen || c.ny > kMaxFftLen || c.nz > kMaxFftLen)
        return fail(nullptr, OHPIC_EINVAL, "axis lengths must be <= %d (one FFT line per LDS block)",
                    kMaxFftLen);
    if (c.weight_order != 0 && c.weight_order != 1)
        return fail(nullptr, OHPIC_EINVAL, "weight_order must be 0 or 1");
    if (c.chunks < 0 || c.chunks > kMaxChunks)
        return fail(nullptr, OHPIC_EINVAL, "chunks must be in [0, %d] (0 = automatic)", kMaxChunks);
    for (int s = 0; s < c.nspecies; ++s) {
        if (c.qm[s] == 0.0) return fail(nullptr, OHPIC_EINVAL, "species %d: qm == 0", s);
        if (c.capacity[s] < 0) return fail(nullptr, OHPIC_EINVAL, "species %d: capacity < 0", s);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(nullptr, OHPIC_ENODEV, "no HIP device available");
    }
    if (c.device < 0 || c.device >= ndev)
        return fail(nullptr, OHPIC_EINVAL, "device %d out of range (%d devices)", c.device, ndev);

    ohpic_ctx *ctx = new ohpic_ctx();
    ctx->cfg = c;
    ctx->device = c.device;
    int rc = OHPIC_OK;
#define CREATE_CHK(expr)                     \
    do {                                     \
        if ((rc = (expr)) != OHPIC_OK) {     \
            std::string e_ = ctx->err;       \
            release(ctx);                    \
            fail(nullptr, rc, "%s", e_.c_str()); \
            return rc;                       \
        }                                    \
    } while (0)
#define CREATE_HIP(expr)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            release(ctx);                                                                   \
            return fail(nullptr, OHPIC_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
        }                                                                                   \
    } while (0)

    CREATE_HIP(hipSetDevice(ctx->device));
    CREATE_HIP(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess)
            ctx->num_cus = prop.multiProcessorCount;
    }
    ctx->n[0] = c.nx;
    ctx->n[1] = c.ny;
    ctx->n[2] = c.nz;
    ctx->ng = (int64_t)c.nx * c.ny * c.nz;
    pick_tile(c, ctx->T);
    if (!tile_supported(ctx->T[0], ctx->T[1], ctx->T[2])) {
        release(ctx);
        return fail(nullptr, OHPIC_EINVAL, "tile %dx%dx%d not compiled in", c.tile[0], c.tile[1],
                    c.tile[2]);
    }
    for (int a = 0; a < 3; ++a) ctx->nt[a] = (ctx->n[a] + ctx->T[a] - 1) / ctx->T[a];
    ctx->ntiles = ctx->nt[0] * ctx->nt[1] * ctx->nt[2];
    ctx->charge_bits = derive_charge_bits(c, ctx->ng);

    InstDev &h = ctx->h;
    for (int a = 0; a < 3; ++a) {
        h.n[a] = ctx->n[a];
        h.T[a] = ctx->T[a];
        h.nt[a] = ctx->nt[a];
        h.b0[a] = c.b0[a];
    }
    h.ntiles = ctx->ntiles;
    h.nspecies = c.nspecies;
    h.weight_order = c.weight_order;
    h.track_ids = c.track_ids ? 1 : 0;
    h.precision = c.precision;
    h.dt = c.dt;
    h.fix_scale = std::ldexp(1.0, ctx->charge_bits);
    h.fix_inv = std::ldexp(1.0, -ctx->charge_bits);
    h.tight_frac = tight_fraction(config_slack(c));
    if (const char *e = std::getenv("OHPIC_TIGHT_LIST_LIMIT")) {
        const long long v = std::atoll(e);
        if (v >= 0 && v < (long long)kTightListCap) ctx->list_limit = (unsigned long long)v;
    }
    const int nt = ctx->ntiles;
    for (int s = 0; s < c.nspecies; ++s) {
        SpeciesHost &S = ctx->sp[s];
        SpecDev &D = h.sp[s];
        S.capacity = c.capacity[s];
        S.alloc = species_alloc(c, s, nt);
        for (int b = 0; b < 2; ++b) {
            if (c.precision == OHPIC_MIXED) {
                for (int k = 0; k < 7; ++k) {
                    uint32_t *p = nullptr;
                    CREATE_CHK(dalloc(ctx, &p, S.alloc));
                    D.buf[b][k] = p;
                }
            } else {
                for (int k = 0; k < 6; ++k) {
                    double *p = nullptr;
                    CREATE_CHK(dalloc(ctx, &p, S.alloc));
                    D.buf[b][k] = p;
                }
                D.buf[b][6] = nullptr;
            }
            D.pid[b] = nullptr;
            if (c.track_ids) CREATE_CHK(dalloc(ctx, &D.pid[b], S.alloc));
            CREATE_CHK(dalloc(ctx, &D.count[b], nt));
            CREATE_HIP(launch_fill_u64(D.count[b], nt, 0ull, ctx->st));
        }
        CREATE_CHK(dalloc(ctx, &D.bin_base, nt));
        CREATE_CHK(dalloc(ctx, &D.bin_cap, nt));
        std::vector<int64_t> zero(nt, 0);
        CREATE_CHK(set_bins(ctx, s, zero));
        D.qm = c.qm[s];
        D.q = c.q_macro[s];
        D.mass = c.q_macro[s] / c.qm[s];
        D.qfix = std::ldexp(c.q_macro[s], ctx->charge_bits);
    }
    for (int b = 0; b < 2; ++b) {
        CREATE_CHK(dalloc(ctx, &h.rho_acc[b], ctx->ng));
        CREATE_HIP(launch_fill_u64(h.rho_acc[b], ctx->ng, 0ull, ctx->st));
    }
    CREATE_CHK(dalloc(ctx, &h.rho, ctx->ng));
    CREATE_CHK(dalloc(ctx, &h.phi, ctx->ng));
    for (int a = 0; a < 3; ++a) CREATE_CHK(dalloc(ctx, &h.e[a], ctx->ng));
    for (double *p : {h.rho, h.phi, h.e[0], h.e[1], h.e[2]})
        CREATE_HIP(launch_fill_f64(p, ctx->ng, 0.0, ctx->st));
    CREATE_CHK(dalloc(ctx, &h.diag, kDiagSlots));
    CREATE_HIP(launch_fill_f64(h.diag, kDiagSlots, 0.0, ctx->st));
    CREATE_CHK(dalloc(ctx, &h.flags, kFlagWords));
    CREATE_CHK(dalloc(ctx, &h.tight_list, kTightListCap));
    CREATE_HIP(launch_fill_u64(h.flags, kFlagWords, 0ull, ctx->st));
    CREATE_CHK(dalloc(ctx, &ctx->d_tmp64, nt));
    CREATE_CHK(dalloc(ctx, &ctx->d_tmp64b, nt));
    CREATE_CHK(dalloc(ctx, &ctx->d_tmp32a, nt));
    CREATE_CHK(dalloc(ctx, &ctx->d_tmp32b, nt));
    CREATE_CHK(dalloc(ctx, &ctx->d_cursor, 1));
    CREATE_CHK(dalloc(ctx, &h.e_partial, 3 * kEfieldMaxBlocks));
    CREATE_CHK(dalloc(ctx, &ctx->d_inst, 1));
    CREATE_CHK(dalloc(ctx, &ctx->d_par01, 2));
    {
        const int p01[2] = {0, 1};
        CREATE_HIP(hipMemcpyAsync(ctx->d_par01, p01, sizeof p01, hipMemcpyHostToDevice, ctx->st));
        CREATE_HIP(hipStreamSynchronize(ctx->st));
    }
    // FFT plans and the Green's-function denominators per axis
    for (int a = 0; a < 3; ++a) {
        CREATE_CHK(make_axis(ctx, ctx->n[a], ctx->ax[a]));
        const int n = ctx->n[a];
        std::vector<double> lam(n);
        const double ref_pi = 3.14159265358979;  // science_constants.f90:4
        for (int k = 0; k < n; ++k) {
            // calc_wave_number, periodic branch (poisson_solver3d.f90:111-117)
            const double wn = (k <= n / 2) ? 2.0 * std::sin(ref_pi * k / (double)n)
                                           : 2.0 * std::sin(ref_pi * (n - k) / (double)n);
            lam[k] = wn * wn;
        }
        CREATE_CHK(dalloc(ctx, &ctx->d_lam[a], n));
        CREATE_HIP(hipMemcpyAsync(ctx->d_lam[a], lam.data(), n * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->st));
        CREATE_HIP(hipStreamSynchronize(ctx->st));
    }
    {
        FftJob &J = ctx->job;
        J.src = h.rho;
        J.data = h.phi;
        for (int a = 0; a < 3; ++a) {
            J.n[a] = ctx->n[a];
            J.ax[a] = ctx->ax[a];
            J.lam[a] = ctx->d_lam[a];
            ctx->fft_grp[a] = fft_groups(a, fft_group_lines(ctx->n[a]), c.nx, c.ny, c.nz);
            ctx->fft_rs[a] = fft_radix_set(ctx->ax[a]);
        }
        J.inv_n = 1.0 / (double)ctx->ng;
        CREATE_CHK(dalloc(ctx, &ctx->d_job, 1));
        CREATE_HIP(hipMemcpyAsync(ctx->d_job, &J, sizeof J, hipMemcpyHostToDevice, ctx->st));
        CREATE_HIP(hipStreamSynchronize(ctx->st));
    }
    CREATE_HIP(hipHostMalloc((void **)&ctx->h_flags, kFlagSlots * sizeof(unsigned long long),
                             hipHostMallocDefault));
    CREATE_HIP(hipHostMalloc((void **)&ctx->h_diag, kDiagSlots * sizeof(double),
                             hipHostMallocDefault));
    CREATE_CHK(sync_upload_inst(ctx));
    CREATE_CHK(build_work_items(ctx));
    CREATE_HIP(hipStreamSynchronize(ctx->st));
#undef CREATE_CHK
#undef CREATE_HIP
    ctx->par = 0;
    ctx->rho_valid = false;
    ctx->binned = true;
    *out = ctx;
    return OHPIC_OK;
}

int ohpic_destroy(ohpic_ctx *ctx)
{
    GateShared gate;
    release(ctx);
    return OHPIC_OK;
}

int ohpic_get_config(const ohpic_ctx *ctx, ohpic_config *out)
{
    if (!ctx || !out) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    *out = ctx->cfg;
    out->tile[0] = ctx->T[0];
    out->tile[1] = ctx->T[1];
    out->tile[2] = ctx->T[2];
    out->charge_bits = ctx->charge_bits;
    return OHPIC_OK;
}

int ohpic_load_particles(ohpic_ctx *ctx, int ispec, int64_t n, const double *x, const double *y,
                         const double *z, const double *vx, const double *vy, const double *vz,
                         const int64_t *pid)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    if (ispec < 0 || ispec >= ctx->cfg.nspecies)
        return fail(ctx, OHPIC_EINVAL, "species %d out of range", ispec);
    if (n < 0 || n > ctx->sp[ispec].capacity)
        return fail(ctx, OHPIC_ENOMEM, "species %d: %lld particles exceed capacity %lld", ispec,
                    (long long)n, (long long)ctx->sp[ispec].capacity);
    if (n > 0 && (!x || !y || !z || !vx || !vy || !vz))
        return fail(ctx, OHPIC_EINVAL, "NULL particle array");
    if (!ctx->binned) return fail(ctx, OHPIC_ESTATE, "particles pushed but not re-binned");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int stg = ctx->par ^ 1;
    SpecDev &D = ctx->h.sp[ispec];
    const double *src[6] = {x, y, z, vx, vy, vz};
    if (ctx->cfg.precision == OHPIC_MIXED) {
        std::vector<int32_t> tile;
        std::vector<uint32_t> word[3];
        std::vector<float> vel[3];
        to_mixed(ctx, n, src, tile, word, vel);
        if (n > 0) {
            HIPCHK(ctx, hipMemcpyAsync(D.buf[stg][0], tile.data(), n * sizeof(int32_t),
                                       hipMemcpyHostToDevice, ctx->st));
            for (int a = 0; a < 3; ++a) {
                HIPCHK(ctx, hipMemcpyAsync(D.buf[stg][1 + a], word[a].data(),
                                           n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->st));
                HIPCHK(ctx, hipMemcpyAsync(D.buf[stg][4 + a], vel[a].data(), n * sizeof(float),
                                           hipMemcpyHostToDevice, ctx->st));
            }
        }
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    } else {
        for (int k = 0; k < 6 && n > 0; ++k)
            HIPCHK(ctx, hipMemcpyAsync(D.buf[stg][k], src[k], n * sizeof(double),
                                       hipMemcpyHostToDevice, ctx->st));
    }
    if (ctx->cfg.track_ids && n > 0) {
        if (pid) {
            HIPCHK(ctx, hipMemcpyAsync(D.pid[stg], pid, n * sizeof(int64_t),
                                       hipMemcpyHostToDevice, ctx->st));
        } else {
            std::vector<int64_t> ids(n);
            std::iota(ids.begin(), ids.end(), (int64_t)0);
            HIPCHK(ctx, hipMemcpyAsync(D.pid[stg], ids.data(), n * sizeof(int64_t),
                                       hipMemcpyHostToDevice, ctx->st));
            HIPCHK(ctx, hipStreamSynchronize(ctx->st));
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    int rc = bin_from_staging(ctx, ispec, n, 0);
    if (rc) return rc;
    ctx->rho_valid = false;
    return build_work_items(ctx);
}

int ohpic_particle_count(ohpic_ctx *ctx, int ispec, int64_t *n)
{
    GateShared gate;
    if (!ctx || !n) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    if (ispec < 0 || ispec >= ctx->cfg.nspecies)
        return fail(ctx, OHPIC_EINVAL, "species %d out of range", ispec);
    *n = ctx->sp[ispec].n;
    return OHPIC_OK;
}

int ohpic_store_particles(ohpic_ctx *ctx, int ispec, int64_t cap, int64_t *n, double *x,
                          double *y, double *z, double *vx, double *vy, double *vz,
                          int64_t *pid)
{
    GateShared gate;
    if (!ctx || !n) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    if (ispec < 0 || ispec >= ctx->cfg.nspecies)
        return fail(ctx, OHPIC_EINVAL, "species %d out of range", ispec);
    if (!ctx->binned) return fail(ctx, OHPIC_ESTATE, "particles pushed but not re-binned");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int64_t cnt = 0;
    int rc = bins_to_staging(ctx, ispec, &cnt);
    if (rc) return rc;
    *n = cnt;
    if (cnt > cap) return fail(ctx, OHPIC_EINVAL, "buffer holds %lld, need %lld", (long long)cap,
                               (long long)cnt);
    const int stg = ctx->par ^ 1;
    SpecDev &D = ctx->h.sp[ispec];
    double *dst[6] = {x, y, z, vx, vy, vz};
    if (ctx->cfg.precision == OHPIC_MIXED) {
        std::vector<int32_t> tile(cnt);
        std::vector<uint32_t> word[3];
        std::vector<float> vel[3];
        if (cnt > 0)
            HIPCHK(ctx, hipMemcpyAsync(tile.data(), D.buf[stg][0], cnt * sizeof(int32_t),
                                       hipMemcpyDeviceToHost, ctx->st));
        for (int a = 0; a < 3; ++a) {
            word[a].resize(cnt);
            vel[a].resize(cnt);
            if (cnt > 0) {
                HIPCHK(ctx, hipMemcpyAsync(word[a].data(), D.buf[stg][1 + a],
                                           cnt * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->st));
                HIPCHK(ctx, hipMemcpyAsync(vel[a].data(), D.buf[stg][4 + a], cnt * sizeof(float),
                                           hipMemcpyDeviceToHost, ctx->st));
            }
        }
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
        from_mixed(ctx, cnt, tile, word, vel, dst);
    } else {
        for (int k = 0; k < 6 && cnt > 0; ++k)
            if (dst[k])
                HIPCHK(ctx, hipMemcpyAsync(dst[k], D.buf[stg][k], cnt * sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->st));
    }
    if (pid && cnt > 0) {
        if (!ctx->cfg.track_ids) return fail(ctx, OHPIC_ESTATE, "ids not tracked");
        HIPCHK(ctx, hipMemcpyAsync(pid, D.pid[stg], cnt * sizeof(int64_t), hipMemcpyDeviceToHost,
                                   ctx->st));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return OHPIC_OK;
}

int ohpic_load_oh_particles(ohpic_ctx *ctx, int64_t n, const ohpic_oh_particle *rec)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    if (n > 0 && !rec) return fail(ctx, OHPIC_EINVAL, "records are NULL");
    const int ns = ctx->cfg.nspecies;
    std::vector<int64_t> cnt(ns, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int s = rec[i].spec - 1;
        if (s < 0 || s >= ns)
            return fail(ctx, OHPIC_EINVAL, "record %lld: spec %d out of range", (long long)i,
                        rec[i].spec);
        cnt[s]++;
    }
    for (int s = 0; s < ns; ++s) {
        std::vector<double> a[6];
        for (auto &v : a) v.reserve(cnt[s]);
        std::vector<int64_t> ids;
        ids.reserve(cnt[s]);
        for (int64_t i = 0; i < n; ++i) {
            if (rec[i].spec - 1 != s) continue;
            a[0].push_back(rec[i].x);
            a[1].push_back(rec[i].y);
            a[2].push_back(rec[i].z);
            a[3].push_back(rec[i].vx);
            a[4].push_back(rec[i].vy);
            a[5].push_back(rec[i].vz);
            ids.push_back(rec[i].pid);
        }
        int rc = ohpic_load_particles(ctx, s, cnt[s], a[0].data(), a[1].data(), a[2].data(),
                                      a[3].data(), a[4].data(), a[5].data(), ids.data());
        if (rc) return rc;
    }
    return OHPIC_OK;
}

int ohpic_store_oh_particles(ohpic_ctx *ctx, int64_t cap, int64_t *n, ohpic_oh_particle *rec)
{
    GateShared gate;
    if (!ctx || !n) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    int64_t tot = 0;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) tot += ctx->sp[s].n;
    *n = tot;
    if (tot > cap) return fail(ctx, OHPIC_EINVAL, "buffer holds %lld, need %lld", (long long)cap,
                               (long long)tot);
    int64_t o = 0;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        const int64_t m = ctx->sp[s].n;
        std::vector<double> a[6];
        for (auto &v : a) v.resize(m);
        std::vector<int64_t> ids(ctx->cfg.track_ids ? m : 0);
        int64_t got = 0;
        int rc = ohpic_store_particles(ctx, s, m, &got, a[0].data(), a[1].data(), a[2].data(),
                                       a[3].data(), a[4].data(), a[5].data(),
                                       ctx->cfg.track_ids ? ids.data() : nullptr);
        if (rc) return rc;
        for (int64_t i = 0; i < got; ++i, ++o) {
            rec[o].x = a[0][i];
            rec[o].y = a[1][i];
            rec[o].z = a[2][i];
            rec[o].vx = a[3][i];
            rec[o].vy = a[4][i];
            rec[o].vz = a[5][i];
            rec[o].pid = ctx->cfg.track_ids ? ids[i] : 0;
            rec[o].nid = 0;
            rec[o].spec = s + 1;
        }
    }
    return OHPIC_OK;
}

int ohpic_seed_uniform_maxwell(ohpic_ctx *ctx, uint64_t initstate, uint64_t initseq,
                               const int64_t *np, const double *vth)
{
    GateShared gate;
    if (!ctx || !np || !vth) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // pcg_setseq_64_srandom_r (pcg_variants.h:790-798) on the host
    Pcg32Dev r0;
    const uint64_t mult = 6364136223846793005ULL;
    r0.state = 0u;
    r0.inc = (initseq << 1u) | 1u;
    r0.state = r0.state * mult + r0.inc;
    r0.state += initstate;
    r0.state = r0.state * mult + r0.inc;
    uint64_t offset = 0;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        if (np[s] < 0 || np[s] > ctx->sp[s].capacity)
            return fail(ctx, OHPIC_ENOMEM, "species %d: %lld particles exceed capacity", s,
                        (long long)np[s]);
        const int stg = ctx->par ^ 1;
        SpecDev &D = ctx->h.sp[s];
        // one random_fix draw, then 7 draws per particle
        if (ctx->cfg.precision == OHPIC_MIXED)
            HIPCHK(ctx, launch_load_uniform_mx(ctx->d_inst, 0, s, stg, r0, offset + 1, np[s],
                                               vth[s], ctx->T, ctx->st));
        else
            HIPCHK(ctx, launch_load_uniform(r0, offset + 1, np[s], ctx->n[0], ctx->n[1],
                                            ctx->n[2], vth[s], (double *)D.buf[stg][0],
                                            (double *)D.buf[stg][1], (double *)D.buf[stg][2],
                                            (double *)D.buf[stg][3], (double *)D.buf[stg][4],
                                            (double *)D.buf[stg][5],
                                            ctx->cfg.track_ids ? D.pid[stg] : nullptr, 0,
                                            ctx->st));
        offset += 1 + 7 * (uint64_t)np[s];
        int rc = bin_from_staging(ctx, s, np[s], 0);
        if (rc) return rc;
    }
    ctx->rho_valid = false;
    return build_work_items(ctx);
}

int ohpic_seed_hump_maxwell(ohpic_ctx *ctx, uint64_t initstate, uint64_t initseq,
                            const int64_t *np, const double *vth, double amp, double sigma,
                            uint64_t budget)
{
    GateShared gate;
    if (!ctx || !np || !vth) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    if (budget < 8) return fail(ctx, OHPIC_EINVAL, "budget must be >= 8 draws");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Pcg32Dev r0;
    const uint64_t mult = 6364136223846793005ULL;
    r0.state = 0u;
    r0.inc = (initseq << 1u) | 1u;
    r0.state = r0.state * mult + r0.inc;
    r0.state += initstate;
    r0.state = r0.state * mult + r0.inc;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        if (np[s] < 0 || np[s] > ctx->sp[s].capacity)
            return fail(ctx, OHPIC_ENOMEM, "species %d: %lld particles exceed capacity", s,
                        (long long)np[s]);
        const int stg = ctx->par ^ 1;
        SpecDev &D = ctx->h.sp[s];
        HIPCHK(ctx, launch_fill_u64(ctx->d_cursor, 1, 0ull, ctx->st));
        if (ctx->cfg.precision == OHPIC_MIXED)
            HIPCHK(ctx, launch_load_hump_mx(ctx->d_inst, 0, s, stg, r0, (uint64_t)s << 40, budget,
                                            np[s], amp, sigma, vth[s], ctx->d_cursor, ctx->T,
                                            ctx->st));
        else
            HIPCHK(ctx, launch_load_hump(r0, (uint64_t)s << 40, budget, np[s], ctx->n[0],
                                         ctx->n[1], ctx->n[2], amp, sigma, vth[s],
                                         (double *)D.buf[stg][0], (double *)D.buf[stg][1],
                                         (double *)D.buf[stg][2], (double *)D.buf[stg][3],
                                         (double *)D.buf[stg][4], (double *)D.buf[stg][5],
                                         ctx->cfg.track_ids ? D.pid[stg] : nullptr, 0,
                                         ctx->d_cursor, ctx->st));
        unsigned long long failed = 0;
        HIPCHK(ctx, hipMemcpyAsync(&failed, ctx->d_cursor, sizeof failed, hipMemcpyDeviceToHost,
                                   ctx->st));
        HIPCHK(ctx, hipStreamSynchronize(ctx->st));
        if (failed)
            return fail(ctx, OHPIC_EINVAL, "rejection sampling exhausted %llu draw budgets",
                        failed);
        int rc = bin_from_staging(ctx, s, np[s], 0);
        if (rc) return rc;
    }
    ctx->rho_valid = false;
    return build_work_items(ctx);
}

int ohpic_deposit(ohpic_ctx *ctx)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    if (!ctx->binned)
        return fail(ctx, OHPIC_ESTATE, "deposit needs binned particles: call boundary first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = do_deposit(ctx);
    if (rc) return rc;
    HIPCHK(ctx, launch_rho_from_acc_b(ctx->d_inst, parity_of(ctx), 1, ctx->ng, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    timers_collect(ctx);
    return OHPIC_OK;
}

int ohpic_field_solve(ohpic_ctx *ctx)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = do_solve(ctx, true);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    timers_collect(ctx);
    return OHPIC_OK;
}

int ohpic_gather_push(ohpic_ctx *ctx)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    if (!ctx->binned)
        return fail(ctx, OHPIC_ESTATE, "gather_push needs binned particles: call boundary");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_particles(kModeGatherPush, ctx->cfg.precision, ctx->cfg.weight_order,
                                 ctx->T, ctx->d_inst, ctx->d_items, ctx->nitems, parity_of(ctx),
                                 ctx->st));
    HIPCHK(ctx, launch_reduce_ke_b(ctx->d_inst, 1, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    ctx->binned = false;
    ctx->rho_valid = false;
    return OHPIC_OK;
}

int ohpic_boundary(ohpic_ctx *ctx)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = zero_next_counts(ctx);
    if (rc) return rc;
    if ((rc = clear_flags(ctx))) return rc;
    HIPCHK(ctx, launch_particles(kModeBoundary, ctx->cfg.precision, ctx->cfg.weight_order, ctx->T,
                                 ctx->d_inst, ctx->d_items, ctx->nitems, parity_of(ctx), ctx->st));
    if ((rc = read_flags(ctx))) return rc;
    timers_collect(ctx);
    rc = absorb_flags(ctx, ctx->h_flags);
    if (rc < 0) return rc;
    if (rc == 1) {
        // a bin overflowed: redo the wrap into a list and bin it into buffer par
        if ((rc = rerun_as_list(ctx, kModeBoundaryList))) return rc;
    } else {
        ctx->par ^= 1;
    }
    ctx->binned = true;
    ctx->rho_valid = false;
    return OHPIC_OK;
}

int ohpic_step(ohpic_ctx *ctx, int nsteps)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    if (nsteps < 0) return fail(ctx, OHPIC_EINVAL, "nsteps < 0");
    if (ctx->failed) return fail(ctx, OHPIC_ESTATE, "an earlier step failed: %s", ctx->err.c_str());
    if (!ctx->binned)
        return fail(ctx, OHPIC_ESTATE, "particles pushed but not re-binned: call boundary");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    for (int i = 0; i < nsteps; ++i) {
        int rc = one_step(ctx);
        if (rc) {
            ctx->failed = true;
            return rc;
        }
    }
    return OHPIC_OK;
}

int ohpic_synchronize(ohpic_ctx *ctx)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    timers_collect(ctx);
    return OHPIC_OK;
}

static int field_ptr(ohpic_ctx *ctx, int which, double **p)
{
    switch (which) {
    case OHPIC_FIELD_RHO: *p = ctx->h.rho; return OHPIC_OK;
    case OHPIC_FIELD_PHI: *p = ctx->h.phi; return OHPIC_OK;
    case OHPIC_FIELD_EX: *p = ctx->h.e[0]; return OHPIC_OK;
    case OHPIC_FIELD_EY: *p = ctx->h.e[1]; return OHPIC_OK;
    case OHPIC_FIELD_EZ: *p = ctx->h.e[2]; return OHPIC_OK;
    default: return fail(ctx, OHPIC_EINVAL, "unknown field %d", which);
    }
}

int ohpic_get_field(ohpic_ctx *ctx, int which, double *values, int nelem, const int32_t lb[3],
                    const int32_t ub[3], int comp)
{
    GateShared gate;
    if (!ctx || !values || !lb || !ub) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    const bool eb = which == OHPIC_FIELD_EB;
    if (nelem < (eb ? 6 : 1) || (!eb && (comp < 0 || comp >= nelem)))
        return fail(ctx, OHPIC_EINVAL, "bad nelem/comp");
    for (int a = 0; a < 3; ++a)
        if (ub[a] < lb[a]) return fail(ctx, OHPIC_EINVAL, "ub < lb");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<double> g[3];
    const int ncomp = eb ? 3 : 1;
    for (int c = 0; c < ncomp; ++c) {
        double *p = nullptr;
        int rc = eb ? field_ptr(ctx, OHPIC_FIELD_EX + c, &p) : field_ptr(ctx, which, &p);
        if (rc) return rc;
        if ((rc = grid_copy_out(ctx, p, g[c]))) return rc;
    }
    const int nx = ctx->n[0], ny = ctx->n[1], nz = ctx->n[2];
    const int64_t sx = ub[0] - lb[0] + 1, sy = ub[1] - lb[1] + 1;
    for (int iz = lb[2]; iz <= ub[2]; ++iz) {
        const int gz = ((iz % nz) + nz) % nz;
        for (int iy = lb[1]; iy <= ub[1]; ++iy) {
            const int gy = ((iy % ny) + ny) % ny;
            for (int ix = lb[0]; ix <= ub[0]; ++ix) {
                const int gx = ((ix % nx) + nx) % nx;
                const int64_t gi = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
                const int64_t hi = (int64_t)nelem * ((ix - lb[0]) + sx * ((iy - lb[1]) +
                                                                          sy * (int64_t)(iz - lb[2])));
                if (eb) {
                    for (int c = 0; c < 3; ++c) values[hi + c] = g[c][gi];
                    for (int c = 0; c < 3; ++c) values[hi + 3 + c] = ctx->cfg.b0[c];
                } else {
                    values[hi + comp] = g[0][gi];
                }
            }
        }
    }
    return OHPIC_OK;
}

int ohpic_put_field(ohpic_ctx *ctx, int which, const double *values, int nelem,
                    const int32_t lb[3], const int32_t ub[3], int comp)
{
    GateShared gate;
    if (!ctx || !values || !lb || !ub) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    if (which == OHPIC_FIELD_EB) return fail(ctx, OHPIC_EINVAL, "put EB: put EX/EY/EZ instead");
    if (nelem < 1 || comp < 0 || comp >= nelem) return fail(ctx, OHPIC_EINVAL, "bad nelem/comp");
    const int nx = ctx->n[0], ny = ctx->n[1], nz = ctx->n[2];
    for (int a = 0; a < 3; ++a)
        if (lb[a] > 0 || ub[a] < ctx->n[a] - 1)
            return fail(ctx, OHPIC_EINVAL, "put: bounds must cover [0, n)");
    double *p = nullptr;
    int rc = field_ptr(ctx, which, &p);
    if (rc) return rc;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<double> g(ctx->ng);
    const int64_t sx = ub[0] - lb[0] + 1, sy = ub[1] - lb[1] + 1;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t hi = (int64_t)nelem * ((ix - lb[0]) + sx * ((iy - lb[1]) +
                                                                          sy * (int64_t)(iz - lb[2])));
                g[ix + (int64_t)nx * (iy + (int64_t)ny * iz)] = values[hi + comp];
            }
    HIPCHK(ctx, hipMemcpyAsync(p, g.data(), ctx->ng * sizeof(double), hipMemcpyHostToDevice,
                               ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return OHPIC_OK;
}

int ohpic_fft3d_r2r(int device, int nx, int ny, int nz, int backward, const double *in,
                    double *out)
{
    GateShared gate;
    if (!in || !out) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(nullptr, OHPIC_EINVAL, "sizes must be >= 1");
    ohpic_ctx tmp;  // holds allocations and the error string only
    tmp.device = device;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, OHPIC_ENODEV, "device %d unavailable", device);
    }
    const int64_t ng = (int64_t)nx * ny * nz;
    int rc = OHPIC_OK;
    FftAxisDev ax[3];
    double *buf = nullptr;
    const int n[3] = {nx, ny, nz};
    hipStream_t st = nullptr;
    auto cleanup = [&]() {
        for (auto &a : tmp.allocs) (void)hipFree(a.first);
        tmp.allocs.clear();
        if (st) (void)hipStreamDestroy(st);
    };
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        cleanup();
        return fail(nullptr, OHPIC_EHIP, "stream creation failed");
    }
    for (int a = 0; a < 3 && !rc; ++a) rc = make_axis(&tmp, n[a], ax[a]);
    if (!rc) rc = dalloc(&tmp, &buf, ng);
    if (rc) {
        std::string e = tmp.err;
        cleanup();
        return fail(nullptr, rc, "%s", e.c_str());
    }
    FftJob job{};
    job.src = buf;
    job.data = buf;
    int64_t grp[3];
    int rs[3];
    for (int a = 0; a < 3; ++a) {
        job.n[a] = n[a];
        job.ax[a] = ax[a];
        grp[a] = fft_groups(a, fft_group_lines(n[a]), nx, ny, nz);
        rs[a] = fft_radix_set(ax[a]);
    }
    job.inv_n = 1.0 / (double)ng;
    FftJob *d_job = nullptr;
    rc = dalloc(&tmp, &d_job, 1);
    if (rc) {
        std::string es = tmp.err;
        cleanup();
        return fail(nullptr, rc, "%s", es.c_str());
    }
    hipError_t e = hipMemcpyAsync(d_job, &job, sizeof job, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(buf, in, ng * sizeof(double), hipMemcpyHostToDevice, st);
    if (!backward) {
        const FftPass p[3] = {kPassR2hcX, kPassR2hcY, kPassR2hcZ};
        for (int a = 0; a < 3 && e == hipSuccess; ++a) e = fft_pass(p[a], d_job, 1, grp[a], rs[a], st);
    } else {
        const FftPass p[3] = {kPassHc2rXScaled, kPassHc2rY, kPassHc2rZ};
        for (int a = 2; a >= 0 && e == hipSuccess; --a) e = fft_pass(p[a], d_job, 1, grp[a], rs[a], st);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, buf, ng * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail(nullptr, OHPIC_EHIP, "fft3d: %s", hipGetErrorString(e));
    return OHPIC_OK;
}

// ------------------------------------------------------- allocation-free FFT
// ohpic_fft3d_r2r_noalloc works in `scratch`: the transform buffer of up to
// OHPIC_FFT_NOALLOC_MAX_POINTS values, the three axes' twiddle tables and the
// job record.  Every GPU the library runs on holds its own copy of the array
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (g_noalloc_ptr).
namespace {
struct NoallocArea {
    double buf[OHPIC_FFT_NOALLOC_MAX_POINTS];
    double2 tw[3][kMaxFftLen];
    FftJob job;
};
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;                          // one call at a time, process-wide
NoallocArea *g_noalloc_ptr[kNoallocMaxDevices];   // scratch's address on each device
double2 g_noalloc_tw[3][kMaxFftLen];              // host copy of the twiddle tables
}  // namespace

static __device__ __align__(256) unsigned char scratch[sizeof(NoallocArea)];

int ohpic_fft3d_r2r_noalloc(int device, int nx, int ny, int nz, int backward, const double *in,
                            double *out)
{
    GateShared gate;
    if (!in || !out) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(nullptr, OHPIC_EINVAL, "sizes must be >= 1");
    const int64_t ng = (int64_t)nx * ny * nz;
    if (ng > OHPIC_FFT_NOALLOC_MAX_POINTS)
        return fail(nullptr, OHPIC_EINVAL,
                    "nx*ny*nz = %lld exceeds the allocation-free maximum of %d points",
                    (long long)ng, OHPIC_FFT_NOALLOC_MAX_POINTS);
    if (device < 0 || device >= kNoallocMaxDevices)
        return fail(nullptr, OHPIC_EINVAL, "device %d out of range", device);
    const int n[3] = {nx, ny, nz};
    FftJob job{};
    int64_t grp[3];
    int rs[3];
    for (int a = 0; a < 3; ++a) {
        FftAxisDev &ax = job.ax[a];
        ax.n = n[a];
        if (!fft_factorize(n[a], ax.fac, &ax.nfac))
            return fail(nullptr, OHPIC_EINVAL,
                        "axis length %d: unsupported (max %d, prime factors <= %d)", n[a],
                        kMaxFftLen, kMaxRadix);
        fft_plan_divisors(ax);
        job.n[a] = n[a];
        grp[a] = fft_groups(a, fft_group_lines(n[a]), nx, ny, nz);
        rs[a] = fft_radix_set(ax);
    }
    job.inv_n = 1.0 / (double)ng;

    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, OHPIC_ENODEV, "device %d unavailable", device);
    }
    if (!g_noalloc_ptr[device]) {
        // first call on this device: its own address of the array
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(nullptr, OHPIC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s",
                        device, hipGetErrorString(e));
        }
        g_noalloc_ptr[device] = static_cast<NoallocArea *>(p);
    }
    NoallocArea *d = g_noalloc_ptr[device];
    for (int a = 0; a < 3; ++a) {
        fft_twiddles(n[a], g_noalloc_tw[a]);
        job.ax[a].tw = d->tw[a];
    }
    job.src = d->buf;
    job.data = d->buf;
    // the null stream; every copy below is complete when the final
    // synchronisation returns, before the lock and the host tables are released
    hipStream_t st = nullptr;
    hipError_t e = hipMemcpyAsync(d->tw, g_noalloc_tw, sizeof g_noalloc_tw,
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&d->job, &job, sizeof job, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d->buf, in, ng * sizeof(double), hipMemcpyHostToDevice, st);
    if (!backward) {
        const FftPass p[3] = {kPassR2hcX, kPassR2hcY, kPassR2hcZ};
        for (int a = 0; a < 3 && e == hipSuccess; ++a) e = fft_pass(p[a], &d->job, 1, grp[a], rs[a], st);
    } else {
        const FftPass p[3] = {kPassHc2rXScaled, kPassHc2rY, kPassHc2rZ};
        for (int a = 2; a >= 0 && e == hipSuccess; --a) e = fft_pass(p[a], &d->job, 1, grp[a], rs[a], st);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d->buf, ng * sizeof(double), hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, OHPIC_EHIP, "fft3d (allocation-free): %s", hipGetErrorString(e));
    }
    return OHPIC_OK;
}

int ohpic_diagnostics(ohpic_ctx *ctx, ohpic_diag *out)
{
    GateShared gate;
    if (!ctx || !out) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_diag, ctx->h.diag, kDiagSlots * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    std::memset(out, 0, sizeof *out);
    out->step = ctx->step;
    double ke = 0.0;
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        out->ke_old[s] = ctx->h_diag[kDiagKeOld + s];
        out->ke_new[s] = ctx->h_diag[kDiagKeNew + s];
        ke += 0.5 * (out->ke_old[s] + out->ke_new[s]);
    }
    out->field_energy = 0.5 * ctx->h_diag[kDiagWe];
    out->field_energy_e2 = 0.5 * ctx->h_diag[kDiagWe2];
    out->net_charge = ctx->h_diag[kDiagQnet];
    out->total_energy = ke + out->field_energy;
    out->repacks = ctx->repacks;
    out->rebalances = ctx->rebalances;
    out->compactions = ctx->compactions;
    out->fast_particles = ctx->fast;
    return OHPIC_OK;
}

int ohpic_energy(ohpic_ctx *ctx, double out[4])
{
    GateShared gate;
    if (!ctx || !out) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    ohpic_diag d;
    const int rc = ohpic_diagnostics(ctx, &d);
    if (rc) return rc;
    out[0] = d.total_energy - d.field_energy;
    out[1] = d.field_energy;
    out[2] = d.total_energy;
    out[3] = d.net_charge;
    return OHPIC_OK;
}

int ohpic_cell_counts(ohpic_ctx *ctx, int ispec, int32_t *counts)
{
    GateShared gate;
    if (!ctx || !counts) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    if (ispec >= ctx->cfg.nspecies) return fail(ctx, OHPIC_EINVAL, "species out of range");
    if (!ctx->binned) return fail(ctx, OHPIC_ESTATE, "particles pushed but not re-binned");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (!ctx->d_cellcounts) {
        int rc = dalloc(ctx, &ctx->d_cellcounts, ctx->ng);
        if (rc) return rc;
    }
    HIPCHK(ctx, launch_fill_i32(ctx->d_cellcounts, ctx->ng, 0, ctx->st));
    for (int s = 0; s < ctx->cfg.nspecies; ++s) {
        if (ispec >= 0 && s != ispec) continue;
        if (ctx->cfg.precision == OHPIC_MIXED)
            HIPCHK(ctx, launch_cell_counts_mx(ctx->d_inst, 0, s, ctx->par, ctx->ntiles,
                                              ctx->d_cellcounts, ctx->T, ctx->st));
        else
            HIPCHK(ctx, launch_cell_counts(ctx->d_inst, 0, s, ctx->par, ctx->ntiles,
                                           ctx->d_cellcounts, ctx->st));
    }
    HIPCHK(ctx, hipMemcpyAsync(counts, ctx->d_cellcounts, ctx->ng * sizeof(int32_t),
                               hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    return OHPIC_OK;
}

int ohpic_enable_timing(ohpic_ctx *ctx, int enable)
{
    GateShared gate;
    if (!ctx) return fail(nullptr, OHPIC_EINVAL, "ctx is NULL");
    timers_collect(ctx);
    ctx->timing = enable != 0;
    for (auto &t : ctx->timers) {
        t.total_ms = 0.0;
        t.launches = 0;
    }
    return OHPIC_OK;
}

int ohpic_kernel_times(ohpic_ctx *ctx, int n, char names[][32], double *total_ms,
                       int64_t *launches, int *count)
{
    GateShared gate;
    if (!ctx || !count) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    HIPCHK(ctx, hipStreamSynchronize(ctx->st));
    timers_collect(ctx);
    int k = 0;
    for (; k < kNumTimers && k < n; ++k) {
        if (names) {
            std::strncpy(names[k], kTimerNames[k], 31);
            names[k][31] = 0;
        }
        if (total_ms) total_ms[k] = ctx->timers[k].total_ms;
        if (launches) launches[k] = ctx->timers[k].launches;
    }
    *count = k;
    return OHPIC_OK;
}

int ohpic_memory_footprint(ohpic_ctx *ctx, int64_t *bytes)
{
    GateShared gate;
    if (!ctx || !bytes) return fail(ctx, OHPIC_EINVAL, "NULL argument");
    *bytes = ctx->bytes;
    return OHPIC_OK;
}

// ------------------------------------------------------------------ groups
int ohpic_group_create(int device, ohpic_group **out)
{
    GateShared gate;
    if (!out) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(nullptr, OHPIC_ENODEV, "no HIP device available");
    }
    if (device < 0 || device >= ndev)
        return fail(nullptr, OHPIC_EINVAL, "device %d out of range (%d devices)", device, ndev);
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, OHPIC_EHIP, "hipSetDevice failed");
    ohpic_group *g = new ohpic_group();
    g->device = device;
    if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return fail(nullptr, OHPIC_EHIP, "stream creation failed");
    }
    if (hipMalloc((void **)&g->d_queue, sizeof(unsigned int)) != hipSuccess ||
        hipMemset(g->d_queue, 0, sizeof(unsigned int)) != hipSuccess) {
        (void)hipGetLastError();
        if (g->d_queue) (void)hipFree(g->d_queue);
        (void)hipStreamDestroy(g->st);
        delete g;
        return fail(nullptr, OHPIC_ENOMEM, "group work counter allocation failed");
    }
    *out = g;
    return OHPIC_OK;
}

int ohpic_group_destroy(ohpic_group *g)
{
    GateShared gate;
    if (!g) return OHPIC_OK;
    (void)hipSetDevice(g->device);
    while (!g->members.empty()) (void)group_detach(g, g->members.back());
    (void)hipStreamSynchronize(g->st);
    drop_graph(g);
    gfree(g->d_insts);
    gfree(g->d_jobs);
    gfree(g->d_parity);
    gfree(g->d_flags);
    gfree(g->d_items);
    gfree(g->d_queue);
    if (g->h_flags) (void)hipHostFree(g->h_flags);
    (void)hipStreamDestroy(g->st);
    delete g;
    return OHPIC_OK;
}

int ohpic_group_attach(ohpic_group *g, ohpic_ctx *ctx)
{
    GateShared gate;
    if (!g || !ctx) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    if (ctx->group) return gfail(g, OHPIC_ESTATE, "context already belongs to a group");
    if (ctx->device != g->device)
        return gfail(g, OHPIC_EINVAL, "context on device %d, group on device %d", ctx->device,
                     g->device);
    if (!ctx->binned) return gfail(g, OHPIC_ESTATE, "particles pushed but not re-binned");
    if (ctx->failed) return gfail(g, OHPIC_ESTATE, "context failed an earlier step");
    if (!g->members.empty()) {
        if (ctx->cfg.precision != g->precision || ctx->cfg.weight_order != g->weight_order ||
            ctx->T[0] != g->T[0] || ctx->T[1] != g->T[1] || ctx->T[2] != g->T[2])
            return gfail(g, OHPIC_EINVAL,
                         "member differs from the group in precision, weight order or tile "
                         "(group %d/%d/%dx%dx%d, context %d/%d/%dx%dx%d)",
                         g->precision, g->weight_order, g->T[0], g->T[1], g->T[2],
                         ctx->cfg.precision, ctx->cfg.weight_order, ctx->T[0], ctx->T[1],
                         ctx->T[2]);
    } else {
        g->precision = ctx->cfg.precision;
        g->weight_order = ctx->cfg.weight_order;
        for (int a = 0; a < 3; ++a) g->T[a] = ctx->T[a];
        g->queue_blocks = 0;
    }
    GHIP(g, hipSetDevice(g->device));
    GHIP(g, hipStreamSynchronize(ctx->st));
    g->members.push_back(ctx);
    g->seen_inst.push_back(0);
    g->seen_items.push_back(0);
    g->dev_par.push_back(-1);
    ctx->group = g;
    g->dirty = true;
    return OHPIC_OK;
}

int ohpic_group_detach(ohpic_group *g, ohpic_ctx *ctx)
{
    GateShared gate;
    if (!g || !ctx) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    return group_detach(g, ctx);
}

int ohpic_group_contains(const ohpic_group *g, const ohpic_ctx *ctx, int *yes)
{
    if (!g || !ctx || !yes) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    *yes = ctx->group == g ? 1 : 0;
    return OHPIC_OK;
}

int ohpic_group_size(const ohpic_group *g, int *n)
{
    if (!g || !n) return fail(nullptr, OHPIC_EINVAL, "NULL argument");
    *n = (int)g->members.size();
    return OHPIC_OK;
}

int ohpic_group_step(ohpic_group *g, int nsteps)
{
    GateShared gate;
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    if (nsteps < 0) return gfail(g, OHPIC_EINVAL, "nsteps < 0");
    return group_step(g, nsteps);
}

int ohpic_group_use_graph(ohpic_group *g, int enable)
{
    GateShared gate;
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    g->use_graph = enable != 0;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->st);
    drop_graph(g);
    return OHPIC_OK;
}

int ohpic_group_set_batch(ohpic_group *g, int steps)
{
    GateShared gate;
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    if (steps < 1 || steps > kGroupMaxBatch)
        return gfail(g, OHPIC_EINVAL, "batch of %d steps: 1 to %d", steps, kGroupMaxBatch);
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->st);
    if (steps != g->batch) {
        g->batch = steps;
        drop_graph(g);
    }
    return OHPIC_OK;
}

int ohpic_group_use_persistent(ohpic_group *g, int enable)
{
    GateShared gate;
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    g->persistent = enable != 0;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->st);
    drop_graph(g);
    return OHPIC_OK;
}

int ohpic_group_stats(const ohpic_group *g, int64_t *steps, int64_t *captures, int64_t *items)
{
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    if (steps) *steps = g->steps;
    if (captures) *captures = g->captures;
    if (items) *items = g->nitems;
    return OHPIC_OK;
}

int ohpic_group_sequences(const ohpic_group *g, int64_t *replayed, int64_t *direct)
{
    if (!g) return fail(nullptr, OHPIC_EINVAL, "group is NULL");
    if (replayed) *replayed = g->replayed;
    if (direct) *direct = g->direct;
    return OHPIC_OK;
}

const char *ohpic_group_last_error(const ohpic_group *g)
{
    if (g) return g->err.c_str();
    return ohpic_last_error(nullptr);
}

// ---------------------------------------------------------------- autotuner
int ohpic_autotune(const ohpic_config *cfgs, int n, int sample, int steps, ohpic_tune_result *out)
{
    GateShared gate;
    if (!cfgs || !out || n <= 0) return fail(nullptr, OHPIC_EINVAL, "bad arguments");
    if (steps < 1) steps = 3;
    if (sample < 1 || sample > n) sample = n;
    std::vector<int> pick(sample);
    for (int k = 0; k < sample; ++k) pick[k] = (int)((int64_t)k * n / sample);
    const int tiles[3][3] = {{4, 4, 4}, {8, 8, 8}, {8, 8, 4}};
    const int chunks[3] = {1, 2, 4};
    std::memset(out, 0, sizeof *out);
    double best = -1.0;
    int timed = 0;
    for (const auto &t : tiles) {
        for (int ch : chunks) {
            ohpic_group *g = nullptr;
            int rc = ohpic_group_create(cfgs[0].device, &g);
            if (rc) return rc;
            std::vector<ohpic_ctx *> ctxs;
            double np_total = 0.0;
            auto cleanup = [&]() {
                ohpic_group_destroy(g);
                for (ohpic_ctx *c : ctxs) ohpic_destroy(c);
            };
            for (int k = 0; k < sample && rc == OHPIC_OK; ++k) {
                ohpic_config c = cfgs[pick[k]];
                for (int a = 0; a < 3; ++a) c.tile[a] = t[a];
                c.chunks = ch;
                ohpic_ctx *ctx = nullptr;
                if ((rc = ohpic_create(&c, &ctx))) break;
                ctxs.push_back(ctx);
                int64_t np[OHPIC_MAX_SPECIES];
                double vth[OHPIC_MAX_SPECIES];
                for (int s = 0; s < c.nspecies; ++s) {
                    np[s] = c.capacity[s];
                    vth[s] = s == 0 ? 1.0 : 0.1;
                    np_total += (double)np[s];
                }
                if ((rc = ohpic_seed_uniform_maxwell(ctx, 42u + (uint64_t)k, 52u, np, vth)))
                    break;
                rc = ohpic_group_attach(g, ctx);
            }
            if (rc == OHPIC_OK) rc = ohpic_group_step(g, 2);
            double rate = 0.0;
            if (rc == OHPIC_OK) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = ohpic_group_step(g, steps);
                const double sec =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                rate = np_total * steps / std::max(sec, 1e-9);
            }
            if (rc != OHPIC_OK) {
                std::string e = ohpic_last_error(nullptr);
                cleanup();
                if (rc == OHPIC_ENOMEM) continue;  // candidate does not fit: skip it
                return fail(nullptr, rc, "autotune %dx%dx%d/%d: %s", t[0], t[1], t[2], ch,
                            e.c_str());
            }
            cleanup();
            timed += 1;
            if (rate > best) {
                best = rate;
                for (int a = 0; a < 3; ++a) out->tile[a] = t[a];
                out->chunks = ch;
                out->particle_steps_per_s = rate;
            }
        }
    }
    out->candidates = timed;
    if (timed == 0) return fail(nullptr, OHPIC_ENOMEM, "autotune: no candidate fits");
    return OHPIC_OK;
}

}  // extern "C"
