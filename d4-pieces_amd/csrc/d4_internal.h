// d4_internal.h -- shared declarations of the host engine and the gfx950
// kernels of the row-by-row tiling search.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "d4.h"

namespace d4 {

constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr int MAX_CAND = 64;       // (piece, transform) pairs of the single-library kernel
constexpr int MAX_LIB_CAND = 128;  // (piece, transform) pairs of any library (two u64 masks)
constexpr int MAX_PIECES = 32;
constexpr int MAX_FIT_ROWS = 6;  // tallest piece
constexpr int MAX_ROWS = 64;     // board height

// Counter slots of the per-step device counter block.
enum : int {
    C32_BAD_ITEM = 0,  // a donated item named a record outside the frontier
    C32_OVERFLOW = 1,  // hash probe limit hit
    C32_EDGES = 2,     // edges written (enum mode)
    C32_EDGE_OVF = 3,  // edge buffer overflow
    C32_OUT = 4,       // compaction output cursor
    C32_DON = 5,       // donated work items written (extend)
    C32_DON_END = 6,   // 1 + end of the valid donated items when C32_DON passed the capacity
    C32_SPILL = 7,     // children that found no slot within the probe limit (count mode)
    C32_DON1 = 8,      // second donation counter pair (follow-up rounds alternate)
    C32_DON_END1 = 9,
    C32_SINK = 10,     // terminal states written to the sink (multi-library batches)
    C32_XOVF = 11,     // a partial row found no room (k_expand spill, k_level buffers)
    C32_LVDONE = 12,   // chunks the level kernels' dedup completed this step
    C32_DONFULL = 13,  // donations / spills refused because the buffer was full
    C32_POVF = 14,     // a bucket or merge pass of the partitioned dedup did not fit
    C32_N = 15,
};
enum : int {
    C64_CHILDREN = 0,
    C64_NEW = 1,
    C64_PRUNED = 2,
    C64_NEXTSIZE = 3,
    C64_WORK = 4,      // dynamic work-item claim (extend)
    C64_DONATED = 5,   // work items handed to follow-up launches
    C64_CHW = 6,       // children written by the level kernels (k_level, incl. holes)
    C64_LVL = 7,       // C64_LVL + L: partial rows written for level L (k_level)
    C64_PSUM_IN = 15,  // partitioned dedup: sum of the surviving children's counts
    C64_PSUM_OUT = 16, // and of the unique states' counts it emitted (equal, mod 2^64)
    C64_LPRUNED = 17,  // children the level kernels pruned (moved to C64_PRUNED by the dedup)
    C64_N = 18,
};

// Arguments of the extend kernel (passed by value).
struct ExtendArgs {
    const uint64_t *in_key;  // records rec_base .. rec_base + n_in - 1
    const uint64_t *in_cnt;
    uint32_t n_in;
    uint32_t rec_base;       // frontier index of in_key[0] (edge parents)
    uint32_t row;            // global step
    uint32_t split;          // levels fixed per work item: 0, 1 or 2
    uint32_t refill;         // idle lanes that trigger a refill round (1..64)
    // library tables (global copies; staged into LDS)
    const uint64_t *fit;     // [8][fit_rows][256] candidate masks
    const uint64_t *placeinfo;  // [8][64][2] placed shape, counter info + range
    const uint64_t *xavail;  // [4][256] candidates ruled out per counter byte
    uint32_t fit_rows;
    // instances
    const uint64_t *bw;      // [n_inst][bw_stride] board window of row r
    uint32_t bw_stride;
    const uint32_t *hgt;     // [n_inst] board heights
    const uint32_t *start;   // [n_inst] step at which the instance joined:
                             // its current row is row - start[inst]
    unsigned long long *result;  // [n_inst] terminal counts
    // key layout
    uint32_t occ_bits;
    uint32_t inst_shift;     // 64 when there is a single instance
    uint64_t occ_mask;
    uint32_t used_mask;
    // dedup table
    unsigned long long *tab_key;
    unsigned long long *tab_cnt;
    uint64_t tab_size;       // slots (any number >= 1)
    uint32_t max_probe;
    // counters
    uint32_t *c32;
    unsigned long long *c64;
    // enum-mode edges
    uint32_t *e_parent;
    uint64_t *e_desc;
    uint64_t *e_slot;
    uint32_t e_cap;
    // pruning
    uint64_t seed_mask;      // window rows no placed piece reaches yet
    uint32_t gcd;
    uint32_t min_size;
    // load balance: a lane whose DFS has run `budget` steps on one work item
    // hands the rest of its search tree to the next launch as donated items
    // {frontier index | first candidate << 32 | level << 40 | only-first <<
    // 48, placements so far}; a launch with don_in set runs such items
    // instead of records
    uint32_t budget;         // 0: never donate
    const uint64_t *don_in;  // [n_don][2] or null
    uint32_t n_don;          // items in don_in, or ~0u: read them from c32[don_read]
    uint32_t don_read;       // counter pair (C32_DON or C32_DON1) that counted don_in
    uint32_t don_write;      // counter pair this launch's donations count into
    uint64_t *don_out;       // [don_cap][2]
    uint32_t don_cap;
    // count mode: a child that finds no slot within the probe limit goes to
    // the spill list (key, count) instead of failing the step; the host then
    // moves the table into a larger one and inserts the spill (0: no spill)
    uint64_t *spill;         // [spill_cap][2]
    uint32_t spill_cap;
};

// Level-synchronous count-mode row step (d4_level.hip).  A partial row
// (state) is {tiling count, occupancy of the rows below cov | the row's open
// cells << 56, use counters | instance << 32}, stored SoA in a buffer of
// `cap` entries per field (open cells 0: an unused slot); its open window
// and candidates are rebuilt from these and the instance's board window when
// it is loaded.  Children are {key, count} SoA.
constexpr uint32_t LVL_FIELDS = 3;
constexpr uint32_t LVL_MAX = 8;  // placements per row (one level each)
// Closed-region areas the connectivity prune allows, as a mask (bit a: at
// least min_size cells and a multiple of g; regions have at most 56 cells).
inline uint64_t area_mask(uint32_t g, uint32_t min_size) {
    uint64_t m = 0;
    for (uint32_t a = 1; a < 64; a++)
        if (a >= min_size && g && a % g == 0) m |= 1ull << a;
    return m;
}
// Control block of row steps replayed from a HIP graph: the kernels of a
// recorded step read the step's sizes here instead of from their launch
// arguments, and the step's last kernel (k_row_close) writes the next step's.
// A step that overflows a buffer sets `fail` and leaves the rest unchanged, so
// the steps queued after it do nothing and its input frontier stays intact.
constexpr uint32_t CTL_LOG = 64;  // steps the host may queue between two looks
struct StepLog {
    unsigned long long n_front, tab_m, children, pruned, fresh, chw, lv_max, out;
};
struct RowCtl {
    unsigned long long n_front;  // records of the next step's frontier
    unsigned long long tab_m;    // its dedup table slots (<= tab_cap)
    unsigned long long tab_cap;  // table slots and next-frontier records allocated
    double ratio;                // unique states per parent of the step before
    uint32_t row;                // its row
    uint32_t fail;               // 1 level buffer, 2 table, 4 spill, 8 sizes,
                                 // 16 a bucket of the partitioned dedup
    uint32_t done;               // steps completed
    uint32_t pad;
    StepLog log[CTL_LOG];        // per completed step, at done % CTL_LOG
};

struct LevelArgs {
    // level 0 reads records; later levels read the states level - 1 wrote
    const uint64_t *in_key, *in_cnt;
    uint32_t n_rec;
    const uint64_t *in_st;   // [LVL_FIELDS][in_cap]
    uint64_t in_cap;
    uint32_t level;          // states read come from c64[C64_LVL + level - 1]
    uint64_t *out_st;        // [LVL_FIELDS][out_cap], counted in c64[C64_LVL + level]
    uint64_t out_cap;
    uint64_t *ch_key, *ch_cnt;  // children, counted in c64[C64_CHW]
    uint64_t ch_cap;
    // library tables (staged into LDS)
    const uint64_t *fit, *placeinfo, *xavail;
    uint32_t fit_rows;
    // instances
    const uint64_t *bw;
    uint32_t bw_stride;
    const uint32_t *hgt, *start;
    unsigned long long *result;
    uint32_t row;
    // key layout
    uint32_t occ_bits, inst_shift, used_mask;
    uint64_t occ_mask;
    uint32_t *c32;
    unsigned long long *c64;
    const RowCtl *ctl;       // set: n_rec and row come from it (graph replay)
    // connectivity prune of each child before it is written (as the dedup's,
    // with its fields); the board windows of row + 1 go to LDS when n_win
    // (the instances) is not 0
    uint32_t prune, n_win;
    // instances whose board window at the row is staged in LDS
    uint32_t n_win0;
    uint64_t area_ok;        // closed-region areas allowed (bit a: area a)
    // with the prune: children gathered per wave in LDS and pruned 64 at a
    // time (else each on its own lane)
    uint32_t stage;
    uint64_t seed_mask;
    uint32_t gcd, min_size, closed_rows;
};
struct DedupArgs {
    const uint64_t *ch_key, *ch_cnt;
    uint64_t ch_cap;         // children: min(c64[C64_CHW], ch_cap)
    const uint64_t *bw;
    uint32_t bw_stride;
    const uint32_t *start;
    uint32_t row;            // the children's row is row + 1
    uint32_t occ_bits, inst_shift;
    uint64_t occ_mask;
    uint64_t seed_mask;
    uint32_t gcd, min_size;
    uint32_t closed_rows;    // max_height - 1: window rows that can hold closed regions
    uint32_t n_inst;         // instances (the children's board windows go to LDS when few)
    uint64_t area_ok;        // closed-region areas allowed (bit a: area a)
    const uint32_t *hgt;
    unsigned long long *tab;  // {key, count} slots
    uint64_t tab_size;
    uint32_t max_probe;
    uint64_t *spill;
    uint32_t spill_cap;
    uint32_t *c32;
    unsigned long long *c64;
    const RowCtl *ctl;       // set: row and tab_size come from it (graph replay)
    // a streamed row step's partition: only children whose key hash has
    // (mix64(key) & part_mask) == part_id are counted and inserted
    uint32_t part_mask, part_id;
};
void launch_level(const LevelArgs &a, int n_cu, hipStream_t st);
void launch_dedup(const DedupArgs &a, bool prune, int n_cu, hipStream_t st);

// Partitioned dedup of a row step's children (d4_part.hip): surviving
// children (connectivity prune) to hash buckets, each bucket merged in one
// workgroup's LDS, unique states written to the next frontier (counted in
// c64[C64_NEW], written in c32[C32_OUT]).  A row with more children than
// n_max <= PART_MAX_CHILDREN sets c32[C32_POVF] and does nothing.
constexpr uint64_t PART_MAX_CHILDREN = 1ull << 30;
struct PartArgs {
    uint64_t *ch_key, *ch_cnt;  // children; after the second radix pass, the
                             // survivors in bucket order
    uint64_t ch_cap;         // children: min(c64[C64_CHW], ch_cap)
    uint64_t n_max;          // children the buffers below are sized for
    uint64_t tile_min;       // children per level-1 tile, at least (part_tile_min())
    uint32_t lp_max;         // log2 of the most buckets (part_lp_max())
    uint32_t slots;          // LDS table slots of the merge (part_slots())
    uint64_t *mask;          // [part_mask_words(n_max)] survivor bits
    uint64_t *pbuf;          // [2 * n_max] {key, count} pairs in coarse-bucket order
    uint64_t *sbuf;          // [2 * n_max] the same pairs in bucket order (second pass)
    uint32_t *hist, *hist2;  // [part_hist_entries(n_max)] per (bucket, tile) counts
    uint32_t *bsum;          // [part_bsum_entries()] scan block sums, then the totals
    uint64_t *out_key, *out_cnt;
    uint64_t out_cap;        // next-frontier records allocated
    // pruning, as DedupArgs
    const uint64_t *bw;
    uint32_t bw_stride;
    const uint32_t *start;
    uint32_t row;            // the children's row is row + 1
    uint32_t inst_shift;
    uint64_t occ_mask;
    uint64_t seed_mask;
    uint32_t gcd, min_size;
    uint32_t closed_rows;
    uint32_t n_inst;
    uint64_t area_ok;        // closed-region areas allowed (bit a: area a)
    uint32_t *c32;
    unsigned long long *c64;
    const RowCtl *ctl;       // set: row comes from it (graph replay)
};
void part_hist_entries(uint64_t n_max, uint64_t tile_min, uint32_t lp_max, uint64_t &level1,
                       uint64_t &level2);
uint64_t part_bsum_entries();
uint64_t part_mask_words(uint64_t n_max);
uint64_t part_tile_min();
uint32_t part_lp_max();
uint32_t part_slots();
// c32[C32_POVF]: bit 0 the buckets did not fit (the step reruns with the
// table), bit 1 the next frontier was too small (C64_NEW says how large it
// must be; launch_part_merge_again reruns the merge alone)
void launch_part_dedup(const PartArgs &a, bool prune, int n_cu, hipStream_t st);
void launch_part_merge_again(const PartArgs &a, int n_cu, hipStream_t st);
size_t level_lds_bytes(uint32_t fit_rows, uint32_t n_win = 0, uint32_t n_win0 = 0,
                       bool stage = false);

// Tables of one library in a multi-library batch (device layout, 16-byte
// aligned).  Candidate masks are 128 bits: word 0 = candidates 0..63.
struct MLib {
    // fit[512 - (512 >> c) + (b >> c)]: candidates anchored at column c whose
    // row-0 cells lie inside row pattern b (b's bits below c are clear)
    uint64_t fit[512][2];
    // fitr[t][c][b]: candidates anchored at column c whose cells in window
    // row t + 1 lie inside row pattern b
    uint64_t fitr[2][8][256][2];
    // xav[b][v]: candidates of the pieces whose use counter lies in byte b
    // of the counters and is exhausted when that byte reads v
    uint64_t xav[4][256][2];
    // per candidate: normalised shape, and info = counter offset |
    // field mask << 8 | count << 16 | anchor column of the shape << 32 |
    // piece's candidate range lo << 40 | hi << 48
    uint64_t cand[MAX_LIB_CAND][2];
    // per piece: offset | field mask << 8 | count << 16 | lo << 32 | hi << 40
    uint64_t piece[MAX_PIECES];
    uint64_t occ_mask;   // occupancy bits of a state
    uint64_t seed_mask;  // child window rows no placed piece reaches
    uint32_t gcd, min_size, n_pieces, n_cand;
};

// What the device needs to build one library's MLib (k_build_mlib): the
// candidates and pieces in MLib's own encoding, plus the candidates' pieces,
// the board width and the pruning constants.
struct MLibSrc {
    uint64_t cand[MAX_LIB_CAND][2];  // as MLib::cand
    uint64_t piece[MAX_PIECES];      // as MLib::piece
    uint8_t cpiece[MAX_LIB_CAND];    // piece of each candidate
    uint32_t W, maxh, gcd, min_size, n_pieces, n_cand;
};
// Builds dst[i] from src[i], i < n, one workgroup per library (d4_tables.hip).
void launch_build_mlib(const MLibSrc *src, MLib *dst, uint32_t n, hipStream_t st);

// Multi-library state key: word 0 = occupancy | instance << MKEY_INST_SHIFT,
// word 1 = use counters.
constexpr uint32_t MKEY_INST_SHIFT = 48;
constexpr uint32_t MAX_MULTI_INST = 1u << 15;

struct MultiArgs {
    const uint64_t *in_k0, *in_k1, *in_cnt;
    uint32_t n_in;
    uint32_t row;            // step = current row of every instance
    uint32_t split;          // levels fixed per work item: 0, 1 or 2
    uint32_t refill;
    const MLib *libs;
    const uint32_t *inst_lib;  // [n_inst] library index
    const uint64_t *bw;        // [n_inst][bw_stride] board window of row r
    uint32_t bw_stride;
    const uint32_t *hgt;
    const uint8_t *dropped;    // [n_inst] nonzero: instance left the search
    unsigned long long *result;
    unsigned long long *tab;   // 4 words per slot: k0, k1, count, unused
    uint64_t tab_size;         // slots (any number; home slot = mulhi(hash, size))
    uint32_t max_probe;
    uint32_t *c32;
    unsigned long long *c64;
    // load balance, as in ExtendArgs (candidate ids up to 127)
    uint32_t budget;
    const uint64_t *don_in;
    uint32_t n_don;          // items in don_in, or ~0u: read them from c32[don_read]
    uint32_t don_read;       // counter pair (C32_DON or C32_DON1) that counted don_in
    uint32_t don_write;      // counter pair this launch's donations count into
    uint64_t *don_out;
    uint32_t don_cap;
    // terminal states of finished instances, two words each: instance |
    // counters << 32, and the state's count (null: not kept)
    uint64_t *sink;
    uint32_t sink_cap;
};

// Launchers (d4_kernels.hip, d4_multi.hip).
// (A non-null ctl, on the kernels of a graph-replayed row step, supplies the
// sizes -- n, m, row -- from device memory, the arguments being the
// capacities, and makes the kernel do nothing after a failed step.)
void launch_fill_table(unsigned long long *key, unsigned long long *cnt,
                       uint64_t n, hipStream_t st, const RowCtl *ctl = nullptr);
void launch_reset_counters(uint32_t *c32, unsigned long long *c64,
                           hipStream_t st, const RowCtl *ctl = nullptr);
// dst[i] = src[i], i < n
void launch_copy_u64(unsigned long long *dst, const unsigned long long *src, uint64_t n,
                     hipStream_t st, const RowCtl *ctl = nullptr);
// a graph-replayed row step's last kernel: checks the step's flags, logs it
// and sets the next step's frontier size, table size and row
void launch_row_close(RowCtl *ctl, const uint32_t *c32, const unsigned long long *c64,
                      hipStream_t st);
void launch_reset_work(unsigned long long *c64, hipStream_t st);
void launch_zero_u32(uint32_t *p, uint64_t n, hipStream_t st);
// blocks x 256 threads x iters random atomic adds into tab[0..m)
void launch_atomic_rate(unsigned long long *tab, uint64_t m, uint32_t iters, uint32_t blocks,
                        hipStream_t st);
// zero the level kernels' per-chunk counters; first: also the overflow flag
// and the completed-chunk count; keep_children: not the children count (the
// chunk's children are appended to the previous chunks')
void launch_reset_levels(uint32_t *c32, unsigned long long *c64, bool first, hipStream_t st,
                         const RowCtl *ctl = nullptr, bool keep_children = false);
// tiling counts of the records whose instance finishes at this step
void launch_terminal(const uint64_t *key, const uint64_t *cnt, uint64_t n, uint32_t row,
                     const uint32_t *start, const uint32_t *hgt, uint32_t inst_shift,
                     unsigned long long *result, int n_cu, hipStream_t st,
                     const RowCtl *ctl = nullptr);
void launch_reset_donations(uint32_t *c32, unsigned long long *c64, hipStream_t st);
// reset one donation counter pair (C32_DON or C32_DON1) and the work counter
void launch_reset_donation_pair(uint32_t *c32, unsigned long long *c64, uint32_t pair,
                                hipStream_t st);
// Move the occupied slots of a {key, count} table of m_old slots into a fresh
// table of m_new slots, then insert n_spill spilled (key, count) records;
// new keys of the spill are counted into c64[C64_NEW], a failed insert sets
// c32[C32_OVERFLOW].
void launch_rehash(const unsigned long long *old_tab, uint64_t m_old, unsigned long long *new_tab,
                   uint64_t m_new, const uint64_t *spill, uint32_t n_spill, uint32_t *c32,
                   unsigned long long *c64, hipStream_t st);
void launch_extend(const ExtendArgs &a, bool enum_mode, bool prune,
                   bool count_only, uint32_t threads, int n_cu, hipStream_t st);
// Count-mode extend over partial-row snapshots (d4_expand.hip): same inputs
// and table semantics as launch_extend with enum_mode = count_only = false.
// In a follow-up launch (don_in set) the items are spilled partial-row
// states of EXPAND_WORDS words each, and don_out receives such states.
constexpr uint32_t EXPAND_WORDS = 5;
void launch_expand(const ExtendArgs &a, bool prune, uint32_t threads, int n_cu, hipStream_t st);
size_t expand_lds_bytes(uint32_t fit_rows, uint32_t threads);
constexpr uint32_t COMPACT_MAX_BLOCKS = 65536;
void launch_compact(const unsigned long long *tab_key,
                    const unsigned long long *tab_cnt, uint64_t m,
                    uint64_t *out_key, uint64_t *out_cnt, uint64_t out_cap, uint32_t *c32,
                    uint32_t *blk, uint32_t *slot_uid, hipStream_t st,
                    const RowCtl *ctl = nullptr);
// multi-library batch: 4-word table slots, two-word keys
void launch_fill_table4(unsigned long long *tab, uint64_t n, hipStream_t st);
void launch_compact4(const unsigned long long *tab, uint64_t m, uint64_t *out_k0,
                     uint64_t *out_k1, uint64_t *out_cnt, uint64_t out_cap, uint32_t *c32,
                     uint32_t *blk, hipStream_t st);
void launch_extend_multi(const MultiArgs &a, bool prune, bool count_only, uint32_t threads,
                         int n_cu, hipStream_t st);
void launch_inst_hist(const uint64_t *k0, uint64_t n, uint32_t n_inst, unsigned long long *hist,
                      int n_cu, hipStream_t st);

// Level-synchronous row step of the multi-library batch (d4_mlevel.hip):
// the single-library level kernels' structure with each row's library read
// through the caches.  A partial row is {tiling count, covered cells below
// >> 8 | the row's open cells << 56, counters | instance << 32} (3 u64, SoA);
// children are {k0, k1, count} SoA, survivors of the prune only (k0 ==
// EMPTY_KEY: an unused slot), counted in c64[C64_CHW].
struct MLevelArgs {
    const uint64_t *in_k0, *in_k1, *in_cnt;  // level 0: the frontier's records
    uint32_t n_rec;
    const uint64_t *in_st;                   // levels >= 1: [LVL_FIELDS][in_cap]
    uint64_t in_cap;
    uint32_t level;
    uint64_t *out_st;                        // [LVL_FIELDS][out_cap]
    uint64_t out_cap;
    uint64_t *ch_k0, *ch_k1, *ch_cnt;
    uint64_t ch_cap;
    const MLib *libs;
    const uint32_t *inst_lib;
    const uint64_t *bw;
    uint32_t bw_stride;
    const uint32_t *hgt;
    const uint8_t *dropped;
    uint32_t row;
    uint32_t prune;
    uint32_t *c32;
    unsigned long long *c64;
};
void launch_mlevel(const MLevelArgs &a, int n_cu, hipStream_t st);
// Terminal records of the frontier (instances finished at this row): their
// counts to the results and the sink (a's table is not used).
void launch_mterminal(const MultiArgs &a, int n_cu, hipStream_t st);
// The children of a completed chunk into a's table (count_only: counted);
// nothing after a level overflow.
void launch_mdedup(const MultiArgs &a, const uint64_t *ch_k0, const uint64_t *ch_k1,
                   const uint64_t *ch_cnt, uint64_t ch_cap, bool count_only, int n_cu,
                   hipStream_t st);
void launch_remap_edges(const uint64_t *e_slot, const uint32_t *slot_uid,
                        uint32_t *e_child, uint32_t n, hipStream_t st);
// out[0..255] = records per current-row pattern, out[256..511] = their
// children's row-0 completion bound
void launch_buckets(const uint64_t *key, uint64_t n, uint32_t row, const uint32_t *start,
                    const uint64_t *bw, uint32_t bw_stride, const uint32_t *hgt,
                    uint32_t inst_shift, uint64_t occ_mask, const unsigned long long *comp,
                    unsigned long long *out, int n_cu, hipStream_t st);
void launch_next_size(const uint64_t *key, uint32_t n, uint32_t row,
                      const uint32_t *start, const uint64_t *bw, uint32_t bw_stride,
                      const uint32_t *hgt, uint32_t inst_shift,
                      uint64_t occ_mask, const unsigned long long *comp,
                      uint32_t *c32, unsigned long long *c64, int n_cu,
                      hipStream_t st);
// first solutions of finished instances (enum mode), see d4_kernels.hip
void launch_first_solutions_row(const uint32_t *par, const uint32_t *chi, const uint64_t *desc,
                                uint64_t n, const uint8_t *live_child, uint8_t *live_par,
                                unsigned long long *best, uint32_t *next, hipStream_t st);
void launch_first_solutions_ends(const uint64_t *keys_term, uint64_t n_term,
                                 const uint64_t *keys_root, uint64_t n_root, uint32_t inst_shift,
                                 const uint8_t *member, uint8_t *live_term, uint32_t *root,
                                 hipStream_t st);
void launch_first_solutions_walk(const uint32_t *slots, uint32_t n_slots, const uint32_t *root,
                                 const uint8_t *live0, const unsigned long long *const *best,
                                 const uint32_t *const *next, uint32_t H, uint64_t *out,
                                 hipStream_t st);
size_t extend_lds_bytes(uint32_t fit_rows, uint32_t threads, bool enum_mode);

// Error plumbing (d4_api.cpp).
void set_error(const std::string &msg);

}  // namespace d4
