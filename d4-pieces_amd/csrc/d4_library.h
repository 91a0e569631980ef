// d4_library.h -- a piece library: counters, (piece, transform) candidates
// and, for the single-library kernel, its LDS tables.
#pragma once

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "d4_internal.h"

using d4::MAX_CAND;
using d4::MAX_LIB_CAND;
using d4::MAX_PIECES;

struct d4_library {
    uint32_t n = 0, sym = 0, W = 0, maxh = 0, ncand = 0, used_bits = 0, fit_rows = 1;
    uint32_t counts[MAX_PIECES] = {}, sizes[MAX_PIECES] = {};
    uint32_t off[MAX_PIECES] = {}, fbits[MAX_PIECES] = {};
    // (piece, transform) candidates in search order, up to MAX_LIB_CAND
    uint8_t cpiece[MAX_LIB_CAND] = {}, ctrs[MAX_LIB_CAND] = {}, cminc[MAX_LIB_CAND] = {};
    uint8_t clo[MAX_LIB_CAND] = {}, chi[MAX_LIB_CAND] = {};  // the piece's candidate range
    uint64_t cnormal[MAX_LIB_CAND] = {};
    uint32_t cinfo[MAX_LIB_CAND] = {};
    uint32_t gcd = 1, min_size = 1;
    uint32_t lds_limit = 65536;
    int device = 0;
    int n_cu = 1;
    uint32_t ext_threads = 1024;  // extend workgroup size (autotuned)
    uint32_t ext_refill = 16;     // idle lanes per refill round (autotuned)
    uint32_t ext_budget = 1024;   // DFS steps per work item before donating (0: never)
    // count-mode row steps: D4_KERNEL_LEVEL (level-synchronous k_level +
    // k_dedup), D4_KERNEL_EXPAND (k_expand: snapshots on a per-wave stack) or
    // D4_KERNEL_DFS (k_extend: one depth-first search per lane)
    uint32_t xp_kernel = D4_KERNEL_LEVEL;
    // count-mode level row steps: D4_STEP_GENERAL / _FUSED / _GRAPH
    uint32_t step_mode = D4_STEP_FUSED;
    // count-mode level row steps: D4_DEDUP_PART (hash buckets merged in LDS)
    // or D4_DEDUP_TABLE (one table in device memory)
    uint32_t dedup_mode = D4_DEDUP_PART;
    uint32_t xp_budget = 256;     // k_expand: tail rounds before a wave's snapshots move on
    uint32_t don_items = 1u << 20;  // donated work items / spilled partial rows per buffer
    // Level-buffer sizes seen by earlier searches on this library, per row
    // step index of a search (widest level and children per parent): the
    // first rows' fan-out is unknown to a search's own history, and a level
    // buffer sized below it reruns its chunk.  Shared by engines on several
    // threads, hence the lock.
    static constexpr uint32_t kSizeRows = 64;
    mutable std::mutex size_mu;
    mutable float seen_lv[kSizeRows] = {}, seen_ch[kSizeRows] = {};
    void note_sizes(uint64_t step, double lv, double ch) const {
        if (step >= kSizeRows) return;
        std::lock_guard<std::mutex> lk(size_mu);
        seen_lv[step] = std::max(seen_lv[step], (float)lv);
        seen_ch[step] = std::max(seen_ch[step], (float)ch);
    }
    void seen_sizes(uint64_t step, double &lv, double &ch) const {
        lv = ch = 0.0;
        if (step >= kSizeRows) return;
        std::lock_guard<std::mutex> lk(size_mu);
        lv = seen_lv[step];
        ch = seen_ch[step];
    }
    // single-library tables (at most MAX_CAND candidates: one u64 mask),
    // built and uploaded on first use by an engine
    std::mutex fast_mu;
    bool fast_ready = false;
    uint64_t place[8][MAX_CAND] = {};
    uint64_t placeinfo[8 * MAX_CAND * 2] = {};
    uint64_t range[MAX_CAND] = {};
    uint64_t prange[MAX_PIECES] = {};
    uint64_t xavail[4 * 256] = {};
    std::vector<uint64_t> fit;  // [8][fit_rows][256]
    unsigned long long comp[256] = {};
    uint64_t *d_fit = nullptr, *d_placeinfo = nullptr, *d_xavail = nullptr;
    unsigned long long *d_comp = nullptr;

    ~d4_library() {
        for (void *p : {(void *)d_fit, (void *)d_placeinfo, (void *)d_xavail, (void *)d_comp})
            if (p) (void)hipFree(p);
    }
};
