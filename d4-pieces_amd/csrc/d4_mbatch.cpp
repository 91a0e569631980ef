// d4_mbatch.cpp -- batches of instances with their own piece libraries and
// board widths: one frontier holds every instance's states (two-word keys
// naming the instance), each row step is one launch of k_extend_multi over
// all of them, and the per-library tables are packed into one device array.
//
// The reference runs one configuration per process: its piece library is a
// process-wide global (g_nme / g_sym, src/frow.h:8-9) and
// SearcherFactory::run iterates board configurations of that one library
// (src/searcher.cpp:86-107).  A batch here mixes libraries and widths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "d4_host.h"
#include "d4_library.h"

namespace d4 {
namespace {

// The candidates and pieces of a library in MLib's encoding, with what the
// device table builder needs besides (k_build_mlib, d4_tables.hip).
void mlib_src(const d4_library &L, MLibSrc &S) {
    std::memset(&S, 0, sizeof S);
    const uint32_t k = L.ncand;
    for (uint32_t j = 0; j < k; j++) {
        S.cand[j][0] = L.cnormal[j];
        S.cand[j][1] = (uint64_t)(L.cinfo[j] & 0xffffu) | ((uint64_t)(L.cinfo[j] >> 16) << 16) |
                       ((uint64_t)L.cminc[j] << 32) | ((uint64_t)L.clo[j] << 40) |
                       ((uint64_t)L.chi[j] << 48);
        S.cpiece[j] = L.cpiece[j];
    }
    for (uint32_t i = 0; i < L.n; i++) {
        uint64_t lo = 0, hi = 0;
        for (uint32_t j = 0; j < k; j++)
            if (L.cpiece[j] == i) {
                lo = L.clo[j];
                hi = L.chi[j];
                break;
            }
        S.piece[i] = (uint64_t)L.off[i] | ((uint64_t)((1u << L.fbits[i]) - 1u) << 8) |
                     ((uint64_t)L.counts[i] << 16) | (lo << 32) | (hi << 40);
    }
    S.W = L.W;
    S.maxh = L.maxh;
    S.gcd = L.gcd;
    S.min_size = L.min_size;
    S.n_pieces = L.n;
    S.n_cand = k;
}

// Device tables of one library built on the host (layout in d4_internal.h).
void build_mlib(const d4_library &L, MLib &M) {
    std::memset(&M, 0, sizeof M);
    const uint32_t k = L.ncand;
    auto placed = [&](uint32_t c, uint32_t j) -> uint64_t {
        const uint64_t n = L.cnormal[j];
        const uint32_t mc = L.cminc[j];
        if (c < mc || c - mc + width8(n) > L.W) return 0;
        return n << (c - mc);
    };
    for (uint32_t c = 0; c < 8; c++) {
        const uint32_t base = 512u - (512u >> c);
        for (uint32_t v = 0; v < (256u >> c); v++) {
            const uint32_t b = v << c;
            uint64_t m[2] = {0, 0};
            for (uint32_t j = 0; j < k; j++) {
                const uint64_t pl = placed(c, j);
                if (!pl || ((uint32_t)(pl & 0xffu) & ~b)) continue;
                m[j >> 6] |= 1ull << (j & 63u);
            }
            M.fit[base + v][0] = m[0];
            M.fit[base + v][1] = m[1];
        }
    }
    // rows 1 and 2 of the window: mark each placement at its own row pattern,
    // then close over supersets (a pattern admits every placement whose
    // cells in that row are a subset of it)
    for (uint32_t t = 0; t < 2; t++)
        for (uint32_t c = 0; c < 8; c++) {
            auto &F = M.fitr[t][c];
            for (uint32_t j = 0; j < k; j++) {
                const uint64_t pl = placed(c, j);
                if (!pl) continue;
                const uint32_t row = (uint32_t)(pl >> (8 * (t + 1))) & 0xffu;
                F[row][j >> 6] |= 1ull << (j & 63u);
            }
            for (uint32_t bit = 0; bit < 8; bit++)
                for (uint32_t b = 0; b < 256; b++)
                    if (b >> bit & 1u) {
                        F[b][0] |= F[b ^ (1u << bit)][0];
                        F[b][1] |= F[b ^ (1u << bit)][1];
                    }
        }
    for (uint32_t j = 0; j < k; j++) {
        M.cand[j][0] = L.cnormal[j];
        M.cand[j][1] = (uint64_t)(L.cinfo[j] & 0xffffu) | ((uint64_t)(L.cinfo[j] >> 16) << 16) |
                       ((uint64_t)L.cminc[j] << 32) | ((uint64_t)L.clo[j] << 40) |
                       ((uint64_t)L.chi[j] << 48);
    }
    for (uint32_t i = 0; i < L.n; i++) {
        uint64_t lo = 0, hi = 0;
        for (uint32_t j = 0; j < k; j++)
            if (L.cpiece[j] == i) {
                lo = L.clo[j];
                hi = L.chi[j];
                break;
            }
        M.piece[i] = (uint64_t)L.off[i] | ((uint64_t)((1u << L.fbits[i]) - 1u) << 8) |
                     ((uint64_t)L.counts[i] << 16) | (lo << 32) | (hi << 40);
    }
    // xav[b][v]: the candidates of every piece whose counter lies in byte b
    // and is exhausted when that byte reads v (a piece of count 0 always is)
    uint64_t pm[MAX_PIECES][2] = {};
    for (uint32_t j = 0; j < k; j++) pm[L.cpiece[j]][j >> 6] |= 1ull << (j & 63u);
    for (uint32_t b = 0; b < 4; b++)
        for (uint32_t v = 0; v < 256; v++)
            for (uint32_t i = 0; i < L.n; i++) {
                if (L.off[i] / 8 != b && !(L.fbits[i] == 0 && b == 0)) continue;
                const uint32_t f = L.fbits[i] ? (v >> (L.off[i] % 8)) & ((1u << L.fbits[i]) - 1u) : 0;
                if (f < L.counts[i]) continue;
                M.xav[b][v][0] |= pm[i][0];
                M.xav[b][v][1] |= pm[i][1];
            }
    const uint32_t occ_bits = 8 * (L.maxh > 1 ? L.maxh - 1 : 0);
    M.occ_mask = occ_bits ? (1ull << occ_bits) - 1ull : 0;
    M.seed_mask = L.maxh >= 2 ? ~((1ull << (8 * (L.maxh - 1))) - 1ull) : ~0ull;
    M.gcd = L.gcd;
    M.min_size = L.min_size;
    M.n_pieces = L.n;
    M.n_cand = k;
}

}  // namespace
}  // namespace d4

using namespace d4;

struct d4_mbatch {
    uint32_t N = 0, flags = 0, max_h = 0, bw_stride = 0, stride = 0;
    std::vector<const d4_library *> lib_of;  // creation inputs (for d4_mbatch_tune)
    std::vector<uint8_t> board_bytes;
    uint32_t threads = 256, refill = 16;
    int device = 0, n_cu = 1;
    std::vector<uint32_t> heights;
    std::vector<unsigned long long> counts;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evx0 = nullptr, evx1 = nullptr, evc0 = nullptr;
    DevBuf<MLib> libs;
    std::vector<uint32_t> lib_index;  // instance -> its library's tables in libs
    DevBuf<uint32_t> inst_lib, hgt;
    DevBuf<uint64_t> bw;
    DevBuf<unsigned long long> result, result_save;
    DevBuf<uint64_t> f0, f1, fc, n0, n1, nc;
    DevBuf<unsigned long long> tab;
    DevBuf<uint32_t> c32, blk;
    DevBuf<unsigned long long> c64;
    DevBuf<uint64_t> don_a, don_b;  // donated work items (two words each)
    // DFS steps per work item of the main launch before its lane hands the
    // rest of its search tree to follow-up launches (0: never)
    uint32_t budget_steps = 4096;
    uint32_t don_items = 1u << 20;  // donated work items per buffer
    uint64_t n_front = 0;
    double grow_unique = 1.0, grow_children = 1.0;
    // unique children per record of the last step and of the step before
    // (0: not measured yet)
    double unique_raw = 0.0, prev_unique_raw = 0.0;
    // bounded memory: frontier halves parked for later (depth first)
    struct Part {
        uint32_t row = 0;
        uint64_t n = 0;
        double grow_unique = 1.0, grow_children = 1.0;
        double unique_raw = 0.0, prev_unique_raw = 0.0;
        DevBuf<uint64_t> k0, k1, c;
    };
    std::vector<std::unique_ptr<Part>> parked;
    uint64_t parked_bytes = 0, mem_cap = 0;
    // per instance: largest frontier share seen, and the state budget beyond
    // which an instance leaves the search unfinished (0 = no budget)
    uint64_t budget = 0;
    std::vector<uint64_t> peak_states;
    std::vector<uint8_t> h_dropped;
    std::vector<unsigned long long> h_hist;
    DevBuf<uint8_t> dropped;
    DevBuf<unsigned long long> hist;
    // terminal states kept for d4_mbatch_terminals: sink_n entries of two
    // words (instance | counters << 32, count) written by finished steps
    bool keep_terminals = false;
    DevBuf<uint64_t> sink;
    uint64_t sink_n = 0, sink_cap = 0, sink_init = 0;  // entries
    std::vector<uint64_t> h_sink;
    static constexpr uint64_t kExactCountMax = 1u << 16;
    static constexpr uint64_t kSampleCount = 1u << 20;  // parents of a sampled count
    bool done = false;
    bool trace = getenv("D4_TRACE") != nullptr;  // per-row lines on stderr
    d4_stats stats{};

    // Copies run on the batch's own (non-blocking) stream, so they are
    // ordered with its kernels (a count restore before the rerun it
    // prepares); the host waits for them.
    void copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) const {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    ~d4_mbatch() {
        if (st) (void)hipStreamSynchronize(st);
        for (hipEvent_t e : {ev0, ev1, evx0, evx1, evc0})
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    void roots() {
        // one root state per instance: nothing placed, counters zero
        std::vector<uint64_t> k0(N), k1(N, 0), c(N, 1);
        for (uint32_t i = 0; i < N; i++) k0[i] = (uint64_t)i << MKEY_INST_SHIFT;
        f0.reserve(N);
        f1.reserve(N);
        fc.reserve(N);
        HIPCHK(hipMemcpyAsync(f0.p, k0.data(), 8ull * N, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(f1.p, k1.data(), 8ull * N, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(fc.p, c.data(), 8ull * N, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(result.p, 0, 8ull * N, st));
        HIPCHK(hipStreamSynchronize(st));
        n_front = N;
        grow_unique = grow_children = 1.0;
        unique_raw = prev_unique_raw = 0.0;
    }

    MultiArgs args(uint32_t row, uint64_t tab_m) const {
        MultiArgs a{};
        a.in_k0 = f0.p;
        a.in_k1 = f1.p;
        a.in_cnt = fc.p;
        a.n_in = (uint32_t)n_front;
        a.row = row;
        a.split = n_front <= 1024 ? 2u : n_front <= (1u << 16) ? 1u : 0u;
        a.refill = refill;
        a.dropped = budget ? dropped.p : nullptr;
        a.libs = libs.p;
        a.inst_lib = inst_lib.p;
        a.bw = bw.p;
        a.bw_stride = bw_stride;
        a.hgt = hgt.p;
        a.result = result.p;
        a.tab = tab.p;
        a.tab_size = tab_m;
        a.max_probe = 256;
        a.c32 = c32.p;
        a.c64 = c64.p;
        a.budget = budget_steps;
        a.don_out = don_a.p;
        a.don_cap = (uint32_t)std::min<uint64_t>(don_a.cap / 2, don_items);
        a.don_read = C32_DON;
        a.don_write = C32_DON;
        if (keep_terminals) {
            a.sink = sink.p + 2 * sink_n;
            a.sink_cap = (uint32_t)std::min<uint64_t>(sink_cap - sink_n, 0xffffffffull);
        }
        return a;
    }

    // One extend pass over the frontier plus the follow-up launches that run
    // the work items lanes over their step budget handed over.
    void extend(uint32_t row, uint64_t tab_m, bool prune, bool count_only) {
        don_a.reserve(2u << 20);
        don_b.reserve(2u << 20);
        launch_extend_multi(args(row, tab_m), prune, count_only, threads, n_cu, st);
        HIPCHK(hipGetLastError());
        if (!budget_steps) return;
        // Rounds are queued a few at a time without a host round trip: each
        // reads its item count from the counter pair the round before it
        // wrote, and the two pairs alternate.  A round lasts about as long
        // as its budget of dependent DFS steps on one lane, so follow-up
        // items get a short budget and split again quickly; after
        // kBudgetedRounds rounds the rest runs to completion.
        constexpr uint32_t kRoundsPerSync = 8;
        constexpr uint32_t kFollowBudget = 64;
        constexpr uint32_t kBudgetedRounds = 96;
        uint32_t rd = C32_DON;
        for (uint32_t round = 0;;) {
            uint32_t h[C32_N];
            copy(h, c32.p, sizeof h, hipMemcpyDeviceToHost);
            if (h[C32_BAD_ITEM]) throw Error(D4_EDEVICE, "internal error: bad donated work item");
            if (h[C32_OVERFLOW]) return;  // the step is rerun
            const uint32_t cap = (uint32_t)std::min<uint64_t>(don_a.cap / 2, don_items);
            const uint32_t n = h[rd] <= cap ? h[rd] : (h[rd + 1] ? h[rd + 1] - 1 : cap);
            if (!n) return;
            for (uint32_t q = 0; q < kRoundsPerSync; q++, round++) {
                const uint32_t wr = rd == C32_DON ? C32_DON1 : C32_DON;
                launch_reset_donation_pair(c32.p, c64.p, wr, st);
                MultiArgs a = args(row, tab_m);
                a.don_in = don_a.p;
                a.n_don = q == 0 ? n : ~0u;
                a.don_read = rd;
                a.don_write = wr;
                a.don_out = don_b.p;
                a.don_cap = (uint32_t)std::min<uint64_t>(don_b.cap / 2, don_items);
                a.budget = round < kBudgetedRounds ? kFollowBudget : 0;
                launch_extend_multi(a, prune, count_only, threads, n_cu, st);
                HIPCHK(hipGetLastError());
                don_a.swap(don_b);
                rd = wr;
            }
        }
    }

    // ---- level kernels (D4_KERNEL_LEVEL, the default) ---------------------
    // The row step as the single-library level kernels run it: one launch
    // per placement level over the frontier (in chunks of at most 8 Mi
    // parents whose buffers fit an eighth of the memory cap), each chunk's
    // surviving children inserted
    // after its levels; instances finishing at this row counted first by a
    // pass of their own.  A chunk whose buffers overflow is rerun alone with
    // four times the room (the chunks before it are complete and inserted).
    bool use_level = true;  // D4_KERNEL_LEVEL (false: D4_KERNEL_DFS)
    DevBuf<uint64_t> lv_a, lv_b, ch0, ch1, chc;
    // partial rows of the widest level, and children, per parent
    double lv_pp = 8.0, ch_pp = 16.0;
    uint64_t level_budget() const { return mem_cap / 8; }
    // bytes per parent of the level buffers (two of LVL_FIELDS words per
    // partial row of the widest level, three words per child, 1.5x room)
    double level_bytes_pp(double lpp, double cpp) const {
        return 8.0 * 1.5 * (2.0 * LVL_FIELDS * lpp + 3.0 * cpp);
    }
    // Parents per chunk: within the budget, and at most 8 Mi (far more rows
    // than the level kernels need to fill the device; larger chunks only
    // make the buffers larger)
    uint64_t level_chunk(uint64_t n, double per) const {
        const uint64_t fit = (uint64_t)((double)level_budget() / per);
        return std::max<uint64_t>(1, std::min<uint64_t>({n, fit, 8ull << 20}));
    }
    // the level buffers a row step of n parents plans with (their floor of
    // 64 Ki entries each is left out: splitting the frontier cannot shrink it)
    uint64_t level_need(uint64_t n) const {
        if (!n) return 0;
        const double per = level_bytes_pp(lv_pp, ch_pp);
        return (uint64_t)(per * (double)level_chunk(n, per));
    }

    // (nf: the first nf parents of the frontier; a count pass may take a
    // prefix, which is a uniform sample since the frontier is in hash order)
    void extend_levels(uint32_t row, uint64_t tab_m, bool prune, bool count_only, uint64_t nf) {
        const MultiArgs m = args(row, tab_m);
        if (!count_only && n_front) launch_mterminal(m, n_cu, st);
        if (!nf) return;
        MLevelArgs a{};
        a.libs = libs.p;
        a.inst_lib = inst_lib.p;
        a.bw = bw.p;
        a.bw_stride = bw_stride;
        a.hgt = hgt.p;
        a.dropped = budget ? dropped.p : nullptr;
        a.row = row;
        a.prune = prune;
        a.c32 = c32.p;
        a.c64 = c64.p;
        double lpp = lv_pp, cpp = ch_pp;
        uint64_t b0 = 0;
        for (uint32_t reruns = 0;;) {
            const double per = level_bytes_pp(lpp, cpp);
            const uint64_t rest = nf - b0;
            const uint64_t chunk = level_chunk(rest, per);
            // (sized exactly, the device pool rounds to 2 MiB; at least 64 Ki
            // entries, room for the waves' partly filled write reservations)
            const uint64_t lcap =
                std::max<uint64_t>((uint64_t)((double)chunk * lpp * 1.5), 1u << 16);
            const uint64_t ccap =
                std::max<uint64_t>((uint64_t)((double)chunk * cpp * 1.5), 1u << 16);
            lv_a.reserve(LVL_FIELDS * lcap);
            lv_b.reserve(LVL_FIELDS * lcap);
            ch0.reserve(ccap);
            ch1.reserve(ccap);
            chc.reserve(ccap);
            a.in_cap = a.out_cap = lcap;
            a.ch_k0 = ch0.p;
            a.ch_k1 = ch1.p;
            a.ch_cnt = chc.p;
            a.ch_cap = ccap;
            const uint32_t nch = (uint32_t)((rest + chunk - 1) / chunk);
            launch_reset_levels(c32.p, c64.p, true, st);
            for (uint32_t j = 0; j < nch; j++) {
                const uint64_t b = b0 + (uint64_t)j * chunk;
                if (j) launch_reset_levels(c32.p, c64.p, false, st);
                a.in_k0 = f0.p + b;
                a.in_k1 = f1.p + b;
                a.in_cnt = fc.p + b;
                a.n_rec = (uint32_t)std::min<uint64_t>(chunk, nf - b);
                for (uint32_t L = 0; L < LVL_MAX; L++) {
                    a.level = L;
                    a.in_st = (L & 1) ? lv_a.p : lv_b.p;
                    a.out_st = (L & 1) ? lv_b.p : lv_a.p;
                    launch_mlevel(a, n_cu, st);
                }
                launch_mdedup(m, ch0.p, ch1.p, chc.p, ccap, count_only, n_cu, st);
            }
            HIPCHK(hipGetLastError());
            uint32_t h32[C32_N];
            unsigned long long h64[C64_N];
            HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (h32[C32_OVERFLOW]) return;  // the table: the step is rerun larger
            const uint32_t done = std::min<uint32_t>(h32[C32_LVDONE], nch);
            const uint64_t c_last = std::min<uint64_t>(chunk, nf - (b0 + (uint64_t)(nch - 1) * chunk));
            if (!h32[C32_XOVF]) {
                // sizes per parent of the last chunk, for the next steps
                if (c_last >= 4096) {
                    uint64_t mx = 0;
                    for (uint32_t L = 0; L < LVL_MAX; L++) mx = std::max<uint64_t>(mx, h64[C64_LVL + L]);
                    lv_pp = std::max(1.0, (double)mx / (double)c_last);
                    ch_pp = std::max(1.0, (double)h64[C64_CHW] / (double)c_last);
                }
                return;
            }
            stats.chunk_reruns++;
            b0 += (uint64_t)done * chunk;
            const uint64_t c_fail = std::min<uint64_t>(chunk, nf - b0);
            bool lv_over = false;
            for (uint32_t L = 0; L < LVL_MAX; L++) lv_over |= h64[C64_LVL + L] > lcap;
            if (lv_over) lpp = 4.0 * std::max(lpp, (double)lcap / (double)c_fail);
            if (h64[C64_CHW] > ccap || !lv_over) cpp = 4.0 * std::max(cpp, (double)ccap / (double)c_fail);
            if (++reruns > 32) throw Error(D4_ERANGE, "a row step's partial rows outgrow every buffer");
        }
    }

    // Children the frontier (the level kernels: its first nf parents) will
    // emit after pruning (search only).
    uint64_t count_children(uint32_t row, uint64_t nf) {
        launch_reset_counters(c32.p, c64.p, st);
        if (use_level) extend_levels(row, 1, !(flags & D4_FLAG_NO_PRUNE), true, nf);
        else extend(row, 1, !(flags & D4_FLAG_NO_PRUNE), true);
        unsigned long long h64[C64_N];
        HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        stats.count_passes++;
        return h64[C64_CHILDREN];
    }

    // Device bytes of a row step with a table of tab_m slots and u children.
    uint64_t need(uint64_t tab_m, uint64_t u) const {
        // (the terminal-state sink, 16 B per state kept, is output and
        // outside the cap: splitting the frontier would not shrink it)
        return 32ull * tab_m + 24ull * u + 24ull * n_front + parked_bytes +
               (use_level ? level_need(n_front) : 0);
    }

    // Park the upper half of the frontier (resumed at this row later).
    void split(uint32_t row) {
        // the failed step's table and next-frontier buffers hold no state:
        // return them to the allocator before the parked half is allocated
        tab.release();
        n0.release();
        n1.release();
        nc.release();
        const uint64_t keep = n_front / 2, m = n_front - keep;
        auto p = std::make_unique<Part>();
        p->row = row;
        p->n = m;
        p->grow_unique = grow_unique;
        p->grow_children = grow_children;
        p->unique_raw = unique_raw;
        p->prev_unique_raw = prev_unique_raw;
        p->k0.reserve(m);
        p->k1.reserve(m);
        p->c.reserve(m);
        HIPCHK(hipMemcpyAsync(p->k0.p, f0.p + keep, 8 * m, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(p->k1.p, f1.p + keep, 8 * m, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(p->c.p, fc.p + keep, 8 * m, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        parked_bytes += 24 * m;
        parked.push_back(std::move(p));
        n_front = keep;
        stats.splits++;
        if (trace)
            fprintf(stderr, "mbatch row %u: split, %llu records parked\n", row,
                    (unsigned long long)m);
    }

    // Resume the most recently parked part; returns its row.
    uint32_t resume() {
        std::unique_ptr<Part> p = std::move(parked.back());
        parked.pop_back();
        parked_bytes -= 24 * p->n;
        f0.swap(p->k0);
        f1.swap(p->k1);
        fc.swap(p->c);
        n_front = p->n;
        grow_unique = p->grow_unique;
        grow_children = p->grow_children;
        unique_raw = p->unique_raw;
        prev_unique_raw = p->prev_unique_raw;
        return p->row;
    }

    // One row step over the whole frontier.  Returns false, with nothing
    // changed, when it would not fit in the memory cap.
    bool row_step(uint32_t row) {
        if (n_front > 0xffffffffull) throw Error(D4_ERANGE, "frontier exceeds 2^32 records");
        const bool prune = !(flags & D4_FLAG_NO_PRUNE);
        // Table size: unique children <= children.  A small frontier gets an
        // exact children count from a search-only pass.  On the level
        // kernels a large one gets the children of its first 1 Mi parents
        // counted (a uniform sample: the frontier is in hash order), scaled;
        // the count pass also sizes the level buffers.  On the depth-first
        // kernel a large one gets an estimate from the previous step's ratios
        // (first rows grow by thousands per record, later ones by a few).
        // Either way the step is rerun at 4x the size if the table still
        // overflows.  (The ratio estimate errs large, up to 20x on the
        // uneven puzzles' early rows: such tables crowd the device, and an
        // allocation that must first release the cached blocks waits
        // seconds for the memory.)
        uint64_t est_u;
        if (n_front <= kExactCountMax) {
            est_u = count_children(row, n_front);
        } else if (use_level) {
            const uint64_t ns = std::min<uint64_t>(n_front, kSampleCount);
            const uint64_t c = count_children(row, ns);
            est_u = (uint64_t)((double)c * ((double)n_front / (double)ns) * 1.1) + 4096;
        } else {
            const uint64_t est_c =
                (uint64_t)((double)n_front * std::min(std::max(grow_children, 1.0), 64.0) * 1.5) +
                4096;
            est_u = std::min<uint64_t>(
                est_c, (uint64_t)((double)n_front * std::max(grow_unique, 0.25) * 1.5) + 4096);
        }
        // (a table has any number of slots: the sampled count sizes it to
        // its estimate; the ratio estimate keeps a power of two's slack)
        uint64_t tab_m = use_level ? round_slots(2 * est_u) : next_pow2(std::max<uint64_t>(2 * est_u, 1024));
        uint32_t h32[C32_N];
        unsigned long long h64[C64_N];
        bool trend = false;
        for (;;) {
            if (n_front >= 2 && (tab_m > (1ull << 33) || need(tab_m, std::min(est_u, tab_m)) > mem_cap)) {
                // A split part cannot merge its equal states with the other
                // part's, so before splitting the step is sized once more
                // from the unique ratio extrapolated from the last two steps
                // (growth slows from row to row); if the table overflows, the
                // rerun at 4x the size does not fit and the split follows.
                if (!trend && n_front > kExactCountMax && prev_unique_raw > unique_raw) {
                    trend = true;
                    const double r = unique_raw * unique_raw / prev_unique_raw;
                    const uint64_t u2 = (uint64_t)((double)n_front * r * 1.5) + 4096;
                    const uint64_t m2 =
                        use_level ? round_slots(2 * u2) : next_pow2(std::max<uint64_t>(2 * u2, 1024));
                    if (m2 < tab_m) {
                        tab_m = m2;
                        est_u = std::min(est_u, u2);
                        continue;
                    }
                }
                return false;
            }
            if (tab_m > (1ull << 33)) throw Error(D4_ERANGE, "dedup table exceeds 2^33 slots");
            try {
                tab.reserve(4 * tab_m);
            } catch (const Error &e) {
                if (e.code != D4_ENOMEM || n_front < 2) throw;
                return false;
            }
            HIPCHK(hipEventRecord(ev0, st));
            HIPCHK(hipMemcpyAsync(result_save.p, result.p, 8ull * N, hipMemcpyDeviceToDevice, st));
            launch_fill_table4(tab.p, tab_m, st);
            launch_reset_counters(c32.p, c64.p, st);
            HIPCHK(hipEventRecord(evx0, st));
            if (n_front) {
                if (use_level) extend_levels(row, tab_m, prune, false, n_front);
                else extend(row, tab_m, prune, false);
                HIPCHK(hipGetLastError());
            }
            HIPCHK(hipEventRecord(evx1, st));
            HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            stats.donation_full += h32[C32_DONFULL];  // (every attempt, reruns included)
            float xms = 0;
            if (n_front) {
                HIPCHK(hipEventElapsedTime(&xms, evx0, evx1));
                stats.extend_ms += xms;
                stats.extend_launches++;
                stats.extend_bytes += 24ull * n_front + 24ull * h64[C64_CHILDREN];
            }
            if (trace)
                fprintf(stderr,
                        "mbatch row %u: n_front %llu table %llu children %llu new %llu "
                        "extend %.3f ms%s\n",
                        row, (unsigned long long)n_front, (unsigned long long)tab_m,
                        (unsigned long long)h64[C64_CHILDREN], (unsigned long long)h64[C64_NEW],
                        xms, h32[C32_OVERFLOW] ? " overflow" : "");
            if (!h32[C32_OVERFLOW]) {
                if (!keep_terminals || sink_n + h32[C32_SINK] <= sink_cap) break;
                // the sink is too small for this step's terminal states:
                // grow it (keeping what earlier steps wrote) and rerun
                grow_sink(sink_n + h32[C32_SINK]);
                stats.retries++;
                copy(result.p, result_save.p, 8ull * N, hipMemcpyDeviceToDevice);
                continue;
            }
            stats.retries++;
            copy(result.p, result_save.p, 8ull * N, hipMemcpyDeviceToDevice);
            tab_m *= 4;
        }
        const uint64_t U = h64[C64_NEW];
        check_row_states(U);
        try {
            n0.reserve(U);
            n1.reserve(U);
            nc.reserve(U);
        } catch (const Error &e) {
            if (e.code != D4_ENOMEM || n_front < 2) throw;
            copy(result.p, result_save.p, 8ull * N, hipMemcpyDeviceToDevice);
            return false;  // (the step's terminal states are written again)
        }
        if (keep_terminals) sink_n += h32[C32_SINK];
        stats.children += h64[C64_CHILDREN];
        stats.pruned += h64[C64_PRUNED];
        stats.donated += h64[C64_DONATED];
        stats.records_extended += n_front;
        stats.unique += U;
        if (n_front) {
            grow_unique = std::max(1.0, (double)U / (double)n_front);
            grow_children = std::max(1.0, (double)h64[C64_CHILDREN] / (double)n_front);
            prev_unique_raw = unique_raw;
            unique_raw = std::max(0.01, (double)U / (double)n_front);
        }
        blk.reserve(2 * COMPACT_MAX_BLOCKS);
        HIPCHK(hipEventRecord(evc0, st));
        launch_compact4(tab.p, tab_m, n0.p, n1.p, nc.p, U, c32.p, blk.p, st);
        uint32_t n_out = 0;
        HIPCHK(hipMemcpyAsync(&n_out, c32.p + C32_OUT, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        check_compacted(n_out, U);
        float ms = 0, cms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&cms, evc0, ev1));
        stats.gpu_ms += ms;
        stats.compact_ms += cms;
        stats.compact_launches++;
        stats.compact_bytes += 8ull * tab_m + 32ull * U;
        f0.swap(n0);
        f1.swap(n1);
        fc.swap(nc);
        n_front = U;
        if (!U) {
            f0.reserve(1);
            f1.reserve(1);
            fc.reserve(1);
        }
        stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, U);
        stats.rows_done++;
        if (U) account(row);
        return true;
    }

    void grow_sink(uint64_t entries) {
        const uint64_t cap = entries + entries / 2 + 1024;
        if (2 * cap > sink.cap) {
            DevBuf<uint64_t> nb;
            nb.reserve(2 * cap);
            if (sink_n)
                HIPCHK(hipMemcpyAsync(nb.p, sink.p, 16ull * sink_n, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
            sink.swap(nb);
        }
        sink_cap = cap;
        stats.sink_growths++;
    }

    // Distinct terminal states of finished instances, sorted, with the
    // counts of parts of a split frontier that reached the same state added.
    void terminals(std::vector<uint64_t> &key, std::vector<uint64_t> &cnt) const {
        key.clear();
        cnt.clear();
        std::vector<std::pair<uint64_t, uint64_t>> v;
        v.reserve(sink_n);
        for (uint64_t k = 0; k < sink_n; k++) {
            const uint64_t e = h_sink[2 * k];
            const uint32_t i = (uint32_t)e;
            if (i < N && !h_dropped[i]) v.emplace_back((e >> 32) | ((uint64_t)i << 32), h_sink[2 * k + 1]);
        }
        std::sort(v.begin(), v.end());
        for (const auto &x : v) {
            if (!key.empty() && key.back() == x.first)
                cnt.back() += x.second;
            else {
                key.push_back(x.first);
                cnt.push_back(x.second);
            }
        }
    }

    // Per-instance frontier shares of the new frontier; instances over the
    // budget are dropped from the next step on.
    void account(uint32_t row) {
        HIPCHK(hipMemsetAsync(hist.p, 0, 8ull * N, st));
        launch_inst_hist(f0.p, n_front, N, hist.p, n_cu, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h_hist.data(), hist.p, 8ull * N, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        bool changed = false;
        for (uint32_t i = 0; i < N; i++) {
            peak_states[i] = std::max<uint64_t>(peak_states[i], h_hist[i]);
            if (budget && h_hist[i] > budget && !h_dropped[i]) {
                h_dropped[i] = 1;
                changed = true;
                if (trace)
                    fprintf(stderr, "mbatch row %u: instance %u dropped (%llu states)\n", row, i,
                            (unsigned long long)h_hist[i]);
            }
        }
        if (changed)
            copy(dropped.p, h_dropped.data(), N, hipMemcpyHostToDevice);
    }

    void run() {
        stats = d4_stats{};
        parked.clear();
        parked_bytes = 0;
        peak_states.assign(N, 0);
        h_dropped.assign(N, 0);
        h_hist.assign(N, 0);
        dropped.reserve(N);
        hist.reserve(N);
        copy(dropped.p, h_dropped.data(), N, hipMemcpyHostToDevice);
        sink_n = 0;
        h_sink.clear();
        if (keep_terminals) {
            sink_cap = sink_init ? sink_init : std::max<uint64_t>(4ull * N, 4096);
            sink.reserve(2 * sink_cap);
        }
        roots();
        // step r extends every instance's row r; an instance of height h is
        // absorbed into its count at step h.  Parts of a split frontier run
        // to the end one after the other; their counts add up per instance.
        uint32_t r = 0;
        for (;;) {
            while (n_front && r <= max_h) {
                if (row_step(r)) {
                    r++;
                    continue;
                }
                if (n_front < 2) throw Error(D4_ENOMEM, "one state's row step exceeds the memory cap");
                split(r);
            }
            if (parked.empty()) break;
            r = resume();
        }
        counts.assign(N, 0);
        copy(counts.data(), result.p, 8ull * N, hipMemcpyDeviceToHost);
        // a dropped instance may have absorbed part of its count before
        // another frontier part crossed the budget: report none of it
        for (uint32_t i = 0; i < N; i++)
            if (h_dropped[i]) counts[i] = 0;
        if (keep_terminals) {
            h_sink.resize(2 * sink_n);
            if (sink_n)
                copy(h_sink.data(), sink.p, 16ull * sink_n, hipMemcpyDeviceToHost);
        }
        done = true;
    }
};

extern "C" {

static std::unique_ptr<d4_mbatch> make_mbatch(uint32_t n_inst, const d4_library *const *libs,
                                              const uint8_t *boards, uint32_t stride,
                                              const uint32_t *heights, uint32_t flags);

int d4_mbatch_create(uint32_t n_inst, const d4_library *const *libs, const uint8_t *boards,
                     uint32_t stride, const uint32_t *heights, uint32_t flags, d4_mbatch **out) {
    return guard([&] {
        if (!out || !libs || !boards || !heights) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        *out = make_mbatch(n_inst, libs, boards, stride, heights, flags).release();
    });
}

// Measured choice of the extend workgroup size and refill threshold on a
// batch of the first `sample` instances (state budget as set on b).
int d4_mbatch_tune(d4_mbatch *b, uint32_t sample, uint32_t *threads, uint32_t *refill,
                   double *best_ms) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        const uint32_t n = std::max(1u, std::min(sample ? sample : 256u, b->N));
        double best = -1;
        uint32_t bt = b->threads, br = b->refill;
        for (uint32_t t : {256u, 512u, 1024u})
            for (uint32_t r : {8u, 16u, 32u}) {
                auto s = make_mbatch(n, b->lib_of.data(), b->board_bytes.data(), b->stride,
                                     b->heights.data(), b->flags);
                s->budget = b->budget;
                s->mem_cap = b->mem_cap;
                s->threads = t;
                s->refill = r;
                s->use_level = false;  // the launch shape is the depth-first kernel's
                s->run();
                const double ms = s->stats.extend_ms;
                if (best < 0 || ms < best) {
                    best = ms;
                    bt = t;
                    br = r;
                }
            }
        b->threads = bt;
        b->refill = br;
        if (threads) *threads = bt;
        if (refill) *refill = br;
        if (best_ms) *best_ms = best;
    });
}

static std::unique_ptr<d4_mbatch> make_mbatch(uint32_t n_inst, const d4_library *const *libs,
                                              const uint8_t *boards, uint32_t stride,
                                              const uint32_t *heights, uint32_t flags) {
    {
        if (!n_inst || n_inst > MAX_MULTI_INST)
            throw Error(D4_EINVAL, "need 1..32768 instances");
        if (flags & ~(uint32_t)(D4_FLAG_NO_PRUNE | D4_FLAG_HOST_TABLES))
            throw Error(D4_EINVAL, "unknown flags");
        auto B = std::make_unique<d4_mbatch>();
        B->N = n_inst;
        B->flags = flags;
        B->heights.assign(heights, heights + n_inst);
        std::map<const d4_library *, uint32_t> index;
        std::vector<const d4_library *> uniq;
        std::vector<uint32_t> il(n_inst);
        for (uint32_t i = 0; i < n_inst; i++) {
            const d4_library *L = libs[i];
            if (!L) throw Error(D4_EINVAL, "null library");
            if (heights[i] > (uint32_t)MAX_ROWS) throw Error(D4_EINVAL, "board taller than 64 rows");
            if (stride < heights[i]) throw Error(D4_EINVAL, "stride shorter than a board");
            const uint32_t occ_bits = 8 * (L->maxh > 1 ? L->maxh - 1 : 0);
            if (occ_bits > MKEY_INST_SHIFT) throw Error(D4_EINVAL, "occupancy exceeds 48 bits");
            const uint8_t wmask = (uint8_t)((1u << L->W) - 1u);
            for (uint32_t r = 0; r < heights[i]; r++)
                if (boards[(size_t)i * stride + r] & ~wmask)
                    throw Error(D4_EINVAL, "board cell outside the library's width");
            auto it = index.find(L);
            if (it == index.end()) {
                if (!uniq.empty() && L->device != uniq[0]->device)
                    throw Error(D4_EINVAL, "libraries on different devices");
                it = index.emplace(L, (uint32_t)uniq.size()).first;
                uniq.push_back(L);
            }
            il[i] = it->second;
            B->max_h = std::max(B->max_h, heights[i]);
        }
        B->device = uniq[0]->device;
        B->n_cu = uniq[0]->n_cu;
        HIPCHK(hipSetDevice(B->device));
        HIPCHK(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking));
        for (hipEvent_t *e : {&B->ev0, &B->ev1, &B->evx0, &B->evx1, &B->evc0})
            HIPCHK(hipEventCreate(e));
        // per-library tables: built on the device from each library's
        // candidates (2.4 KB sent per library instead of 92 KB of tables),
        // or on the host (D4_FLAG_HOST_TABLES)
        B->libs.reserve(uniq.size());
        std::vector<MLib> ml;
        std::vector<MLibSrc> ms;
        DevBuf<MLibSrc> msrc;
        if (flags & D4_FLAG_HOST_TABLES) {
            ml.resize(uniq.size());
            for (size_t j = 0; j < uniq.size(); j++) build_mlib(*uniq[j], ml[j]);
            HIPCHK(hipMemcpyAsync(B->libs.p, ml.data(), sizeof(MLib) * ml.size(),
                                  hipMemcpyHostToDevice, B->st));
        } else {
            ms.resize(uniq.size());
            for (size_t j = 0; j < uniq.size(); j++) mlib_src(*uniq[j], ms[j]);
            msrc.reserve(ms.size());
            HIPCHK(hipMemcpyAsync(msrc.p, ms.data(), sizeof(MLibSrc) * ms.size(),
                                  hipMemcpyHostToDevice, B->st));
            launch_build_mlib(msrc.p, B->libs.p, (uint32_t)ms.size(), B->st);
            HIPCHK(hipGetLastError());
        }
        B->lib_index = il;
        B->inst_lib.reserve(n_inst);
        HIPCHK(hipMemcpyAsync(B->inst_lib.p, il.data(), 4ull * n_inst, hipMemcpyHostToDevice,
                              B->st));
        B->hgt.reserve(n_inst);
        HIPCHK(hipMemcpyAsync(B->hgt.p, heights, 4ull * n_inst, hipMemcpyHostToDevice, B->st));
        // board windows: bw[i][r] = rows r .. r+7 of instance i (0 past the
        // bottom), for r = 0 .. max_h + 1
        B->bw_stride = B->max_h + 2;
        std::vector<uint64_t> w((size_t)n_inst * B->bw_stride, 0);
        for (uint32_t i = 0; i < n_inst; i++)
            for (uint32_t r = 0; r < B->bw_stride; r++) {
                uint64_t v = 0;
                for (uint32_t j = 0; j < 8; j++)
                    if (r + j < heights[i])
                        v |= (uint64_t)boards[(size_t)i * stride + r + j] << (8 * j);
                w[(size_t)i * B->bw_stride + r] = v;
            }
        B->bw.reserve(w.size());
        HIPCHK(hipMemcpyAsync(B->bw.p, w.data(), 8 * w.size(), hipMemcpyHostToDevice, B->st));
        B->result.reserve(n_inst);
        B->result_save.reserve(n_inst);
        B->c32.reserve(C32_N);
        B->c64.reserve(C64_N);
        HIPCHK(hipStreamSynchronize(B->st));
        B->mem_cap = (uint64_t)(0.8 * (double)available_device_bytes());
        B->lib_of.assign(libs, libs + n_inst);
        B->board_bytes.assign(boards, boards + (size_t)n_inst * stride);
        B->stride = stride;
        return B;
    }
}

int d4_mbatch_set_launch(d4_mbatch *b, uint32_t threads, uint32_t refill) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (threads != 256 && threads != 512 && threads != 1024)
            throw Error(D4_EINVAL, "threads must be 256, 512 or 1024");
        if (refill < 1 || refill > 64) throw Error(D4_EINVAL, "refill must be 1..64");
        b->threads = threads;
        b->refill = refill;
    });
}

int d4_mbatch_set_budget_steps(d4_mbatch *b, uint32_t steps) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        b->budget_steps = steps;
    });
}

int d4_mbatch_set_donation_capacity(d4_mbatch *b, uint32_t items) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (items < 64) throw Error(D4_EINVAL, "donation capacity below 64 items");
        b->don_items = items;
    });
}

int d4_mbatch_set_kernel(d4_mbatch *b, uint32_t kernel) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (kernel != D4_KERNEL_LEVEL && kernel != D4_KERNEL_DFS)
            throw Error(D4_EINVAL, "a multi-library batch runs the level or the depth-first kernel");
        b->use_level = kernel == D4_KERNEL_LEVEL;
    });
}

int d4_mbatch_set_memory(d4_mbatch *b, uint64_t bytes) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (bytes < (1u << 20)) throw Error(D4_EINVAL, "memory cap below 1 MiB");
        b->mem_cap = bytes;
    });
}

int d4_mbatch_set_budget(d4_mbatch *b, uint64_t max_states) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        b->budget = max_states;
    });
}

int d4_mbatch_instances(const d4_mbatch *b, uint64_t *peak_states, uint8_t *finished) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (!b->done) throw Error(D4_ESTATE, "batch not run");
        for (uint32_t i = 0; i < b->N; i++) {
            if (peak_states) peak_states[i] = b->peak_states[i];
            if (finished) finished[i] = b->h_dropped[i] ? 0 : 1;
        }
    });
}

int d4_mbatch_run(d4_mbatch *b) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        HIPCHK(hipSetDevice(b->device));
        b->run();
    });
}

int d4_mbatch_counts(const d4_mbatch *b, uint64_t *counts) {
    return guard([&] {
        if (!b || !counts) throw Error(D4_EINVAL, "null argument");
        if (!b->done) throw Error(D4_ESTATE, "batch not run");
        for (uint32_t i = 0; i < b->N; i++) counts[i] = b->counts[i];
    });
}

int d4_mbatch_stats(const d4_mbatch *b, d4_stats *st) {
    return guard([&] {
        if (!b || !st) throw Error(D4_EINVAL, "null argument");
        *st = b->stats;
    });
}

int d4_mbatch_set_terminals(d4_mbatch *b, int keep, uint64_t initial_entries) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null argument");
        b->keep_terminals = keep != 0;
        b->sink_init = initial_entries;
    });
}

int d4_mbatch_terminals(const d4_mbatch *b, uint32_t *inst, uint64_t *counters, uint64_t *tilings,
                        size_t cap, size_t *n) {
    return guard([&] {
        if (!b || !n) throw Error(D4_EINVAL, "null argument");
        if (!b->done || !b->keep_terminals)
            throw Error(D4_ESTATE, "run the batch with terminal states kept first");
        std::vector<uint64_t> key, cnt;
        b->terminals(key, cnt);
        *n = key.size();
        for (size_t k = 0; k < key.size() && k < cap; k++) {
            if (inst) inst[k] = (uint32_t)(key[k] >> 32);
            if (counters) counters[k] = key[k] & 0xffffffffull;
            if (tilings) tilings[k] = cnt[k];
        }
    });
}

void d4_mbatch_destroy(d4_mbatch *b) { delete b; }

}  // extern "C"

uint64_t d4_mlib_table_bytes(void) { return sizeof(MLib); }

int d4_mbatch_tables(const d4_mbatch *b, uint32_t inst, void *out, uint64_t bytes) {
    return guard([&] {
        if (!b || !out) throw Error(D4_EINVAL, "null argument");
        if (inst >= b->N) throw Error(D4_EINVAL, "instance out of range");
        if (bytes < sizeof(MLib)) throw Error(D4_ERANGE, "buffer smaller than the tables");
        HIPCHK(hipSetDevice(b->device));
        HIPCHK(hipMemcpyAsync(out, b->libs.p + b->lib_index[inst], sizeof(MLib),
                              hipMemcpyDeviceToHost, b->st));
        HIPCHK(hipStreamSynchronize(b->st));
    });
}

int d4_library_host_tables(const d4_library *lib, void *out, uint64_t bytes) {
    return guard([&] {
        if (!lib || !out) throw Error(D4_EINVAL, "null argument");
        if (bytes < sizeof(MLib)) throw Error(D4_ERANGE, "buffer smaller than the tables");
        auto m = std::make_unique<MLib>();
        build_mlib(*lib, *m);
        std::memcpy(out, m.get(), sizeof(MLib));
    });
}
