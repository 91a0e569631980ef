// d4_level.hip -- level-synchronous count-mode row step (gfx950).
//
// A row's completions form a tree: level L holds the partial rows with L
// placements (each placement covers the row's lowest open cell, the
// reference CPU searcher's order, src/Piece.cpp:82-98), and a row has at
// most 8 levels.  Instead of one depth-first search per lane (k_extend) the
// row step runs one launch per level over every partial row of the frontier:
//
//   k_level   a wave loads 64 partial rows (level 0: records, expanded to
//             their root partial row), lists their (row, candidate) pairs in
//             LDS (a wave prefix sum of the candidate counts gives each row
//             its range of the list) and then processes the pairs 64 at a
//             time, one per lane, so no lane idles on a row with few
//             candidates or waits on a sibling's deeper subtree.  A placement
//             that completes the row yields a child {key, count}; one that
//             leaves it open yields a partial row of the next level (24 B:
//             its window and candidates are rebuilt when it is loaded) -- or
//             nothing, when it has no candidate.  Completed rows are gathered
//             in the wave's LDS staging area and pruned 64 at a time (the
//             connectivity test on every lane), and only the survivors are
//             written.  Outputs are written densely (ballot + mbcnt) into
//             chunks that each wave reserves with one atomic per chunk, not
//             per record.
//   k_dedup   merges each child into the open-addressing table (CAS, counts
//             summed; pruning it first when the level kernels did not), with
//             nothing else on its lanes' path, so many probes are in flight
//             at once.  Large rows use the partitioned dedup (d4_part.hip).
//
// The surviving children are the ones k_extend keeps (each child is one
// placement sequence; only the order of discovery differs).  The
// reference's counterparts are tiled_row_search / impl<H>
// (src/kernel.cuh:79-160) and the host Sorter (src/sorter.cpp:156-214).
#include "d4_internal.h"
#include "d4_device.h"

#include <algorithm>

namespace d4 {

namespace {

using namespace dev;

constexpr int LT = 1024;          // threads per workgroup (16 waves: 4 per SIMD)
constexpr uint32_t PCAP = 1024;   // listed (row, candidate) pairs per wave and pass
constexpr uint32_t OWN = 5;       // per-wave LDS words per loaded row: E, cov, cnt, ui, avail

// The board window of row + 1 of instance i (0 past the instance's rows,
// as the dedup stages it).
__device__ __forceinline__ uint64_t child_window(const LevelArgs &a, uint32_t i) {
    const uint32_t rr = a.row + 1 - a.start[i];
    return rr < a.bw_stride ? a.bw[(size_t)i * a.bw_stride + rr] : 0ull;
}

// Connectivity prune of the child with key bits kb (key without the
// instance) of instance inst, as the dedup's (k_dedup below): false, and
// one more pruned, when the open cells it leaves hold a closed region that
// no set of pieces can fill.
__device__ __forceinline__ bool child_ok(const LevelArgs &a, const uint64_t *s_win, uint32_t inst,
                                         uint64_t kb, uint64_t &pruned) {
    const uint64_t w = inst < a.n_win ? s_win[inst] : child_window(a, inst);
    const uint64_t E = w & ~(kb & a.occ_mask);
    const bool ok = a.closed_rows <= 4 ? viable32(E, a.closed_rows, AreaMask{a.area_ok})
                                       : viable(E, a.seed_mask, AreaMask{a.area_ok});
    pruned += !ok;
    return ok;
}

template <bool FROM_REC, int FR>
__global__ __launch_bounds__(LT) void k_level(LevelArgs a) {
    extern __shared__ uint64_t lds[];
    if (a.ctl) {
        // graph replay: the launch carries capacities, the step's sizes are here
        if (a.ctl->fail) return;
        a.n_rec = (uint32_t)min((uint64_t)a.ctl->n_front, (uint64_t)a.n_rec);
        a.row = a.ctl->row;
    }
    // the level's rows, written by the launch before (every thread reads the
    // same count): a workgroup with no batch of them -- its first row, the
    // first of its first wave, lies past the level's end -- leaves before
    // staging the tables
    uint64_t n_in;
    if (FROM_REC) {
        n_in = a.n_rec;
    } else {
        n_in = ctr_load(a.c64 + C64_LVL + a.level - 1);
        if (n_in > a.in_cap) n_in = a.in_cap;
    }
    if ((uint64_t)blockIdx.x * LT >= n_in) return;
    const uint32_t nfit = 8u * a.fit_rows * 256u;
    uint64_t *s_fit = lds;
    const PlaceInfo *s_pc = reinterpret_cast<const PlaceInfo *>(s_fit + nfit);
    uint64_t *s_xa = s_fit + nfit + 2 * 8 * MAX_CAND;
    uint64_t *s_rng = s_xa + 4 * 256;
    // (after the waves' own areas: the board windows of row + 1 for the
    // prune, then each instance's window at the row and whether the row is
    // one of its rows)
    uint64_t *s_win = s_rng + MAX_CAND + (size_t)(LT / 64) * (OWN * 64 + PCAP / 4);
    uint64_t *s_w0 = s_win + (a.prune ? a.n_win : 0u);
    const uint32_t nw0 = a.n_win0;
    uint8_t *s_t0 = reinterpret_cast<uint8_t *>(s_w0 + nw0);
    // (then, when the prune runs on gathered children, each wave's staging
    // area: 64 child keys and their 64 counts)
    const bool stage = a.prune && a.stage;
    uint64_t *s_stg = s_w0 + nw0 + (nw0 + 7) / 8;
    for (uint32_t i = threadIdx.x; i < nfit; i += LT) s_fit[i] = a.fit[i];
    for (uint32_t i = threadIdx.x; i < 2 * 8 * MAX_CAND; i += LT) s_fit[nfit + i] = a.placeinfo[i];
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += LT) s_xa[i] = a.xavail[i];
    for (uint32_t i = threadIdx.x; i < MAX_CAND; i += LT)
        s_rng[i] = piece_range(a.placeinfo[2 * i + 1]);
    if (a.prune)
        for (uint32_t i = threadIdx.x; i < a.n_win; i += LT) s_win[i] = child_window(a, i);
    for (uint32_t i = threadIdx.x; i < nw0; i += LT) {
        const uint32_t rr = a.row - a.start[i];
        const bool in = rr < a.hgt[i];
        s_w0[i] = in ? a.bw[(size_t)i * a.bw_stride + rr] : 0ull;
        s_t0[i] = in;
    }
    __syncthreads();

    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    uint64_t *o_E = s_rng + MAX_CAND + (size_t)wv * (OWN * 64 + PCAP / 4);
    uint64_t *o_cov = o_E + 64, *o_cnt = o_cov + 64, *o_ui = o_cnt + 64, *o_av = o_ui + 64;
    uint16_t *plist = reinterpret_cast<uint16_t *>(o_av + 64);
    uint64_t *st_key = s_stg + (size_t)wv * 128, *st_cnt = st_key + 64;

    if (__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint64_t waves = (uint64_t)gridDim.x * (LT / 64);
    const uint64_t wid = (uint64_t)blockIdx.x * (LT / 64) + wv;
    // chunk size: small levels get small chunks (fewer holes)
    uint64_t chunk = n_in / (waves ? waves : 1);
    chunk = chunk < 64 ? 64 : chunk > 1024 ? 1024 : chunk;
    Writer ws, wc;
    bool alive = true;
    uint64_t pruned = 0;
    unsigned long long *ctr_s = a.c64 + C64_LVL + a.level, *ctr_c = a.c64 + C64_CHW;
    uint64_t *oCnt = a.out_st, *oCw = oCnt + a.out_cap, *oUi = oCw + a.out_cap;
    // the board window of instance i at the row
    auto row_window = [&](uint32_t i) -> uint64_t {
        return i < nw0 ? s_w0[i] : a.bw[(size_t)i * a.bw_stride + (a.row - a.start[i])];
    };

    // Children to the children buffer: lanes with e hold one {key, count}.
    // With the prune, either each child is tested on its own lane before it
    // is written, or (stage) the wave's children are gathered in its
    // staging area and tested 64 at a time, so that the flood fill runs on
    // every lane.  False when the buffer is full.
    uint32_t n_s = 0;  // children staged (wave-uniform)
    auto write_children = [&](bool keep, uint64_t key, uint64_t cnt) -> bool {
        const uint64_t km = __ballot(keep);
        if (!km) return true;
        bool okw = true;
        const uint64_t sl = place(wc, km, ctr_c, chunk, a.ch_cap, a.c32, okw);
        if (!okw) return false;
        if (keep) {
            a.ch_key[sl] = key;
            a.ch_cnt[sl] = cnt;
        }
        return true;
    };
    auto inst_of = [&](uint64_t key) -> uint32_t {
        return a.inst_shift < 64 ? (uint32_t)(key >> a.inst_shift) : 0u;
    };
    auto flush_staged = [&](uint32_t n) -> bool {
        __builtin_amdgcn_wave_barrier();
        bool keep = false;
        uint64_t key = 0, cnt = 0;
        if (lane < n) {
            key = st_key[lane];
            cnt = st_cnt[lane];
            keep = child_ok(a, s_win, inst_of(key), key, pruned);
        }
        __builtin_amdgcn_wave_barrier();
        return write_children(keep, key, cnt);
    };
    auto emit_children = [&](bool e, uint64_t key, uint64_t cnt) -> bool {
        if (!stage) return write_children(e && (!a.prune || child_ok(a, s_win, inst_of(key), key, pruned)),
                                          key, cnt);
        const uint64_t m = __ballot(e);
        if (!m) return true;
        const uint32_t nn = (uint32_t)__builtin_popcountll(m);
        const uint32_t slot = n_s + mbcnt(m);
        __builtin_amdgcn_wave_barrier();
        if (e && slot < 64) {
            st_key[slot] = key;
            st_cnt[slot] = cnt;
        }
        if (n_s + nn < 64) {
            n_s += nn;
            return true;
        }
        if (!flush_staged(64)) return false;
        if (e && slot >= 64) {
            st_key[slot - 64] = key;
            st_cnt[slot - 64] = cnt;
        }
        n_s = n_s + nn - 64;
        return true;
    };

    // the next batch's inputs are loaded while the current one is processed
    const uint64_t stride = waves * 64;
    uint64_t nx0 = 0, nx1 = 0, nx2 = 0;
    auto fetch = [&](uint64_t b) {
        const uint64_t i = b + lane;
        if (i < n_in) {
            if (FROM_REC) {
                nx0 = a.in_key[i];
                nx1 = a.in_cnt[i];
            } else {
                const uint64_t *s = a.in_st;
                nx0 = s[i];
                nx1 = s[a.in_cap + i];
                nx2 = s[2 * a.in_cap + i];
            }
        }
    };
    if (wid * 64 < n_in) fetch(wid * 64);
    for (uint64_t b0 = wid * 64; alive && b0 < n_in; b0 += stride) {
        const uint64_t i = b0 + lane;
        const bool valid = i < n_in;
        const uint64_t c0 = nx0, c1 = nx1, c2 = nx2;
        if (b0 + stride < n_in) fetch(b0 + stride);
        uint64_t E = 0, ok = 0, cnt = 0, cov = 0, av = 0;
        uint32_t used = 0, inst = 0;
        bool pass = false;
        if (FROM_REC) {
            if (valid) {
                const uint64_t key = c0;
                cnt = c1;
                inst = a.inst_shift < 64 ? (uint32_t)(key >> a.inst_shift) : 0u;
                used = (uint32_t)(key >> a.occ_bits) & a.used_mask;
                // a terminal state (its instance finishes at this step) is
                // counted by the step's terminal pass and has no children
                bool in;
                uint64_t w;
                if (inst < nw0) {
                    in = s_t0[inst] != 0;
                    w = s_w0[inst];
                } else {
                    const uint32_t rr = a.row - a.start[inst];
                    in = rr < a.hgt[inst];
                    w = in ? a.bw[(size_t)inst * a.bw_stride + rr] : 0ull;
                }
                if (in) {
                    const uint64_t occ = key & a.occ_mask;
                    E = w & ~occ;
                    cov = occ >> 8;
                    if (E & 0xffu) {
                        const uint32_t c = (uint32_t)__builtin_ctzll(E & 0xffu);
                        av = avail_of(s_xa, used);
                        ok = fitmask<FR>(s_fit, a.fit_rows, c, E) & av;
                    } else {
                        pass = true;  // row already covered from above: its own child
                    }
                }
            }
        } else if (valid) {
            cnt = c0;
            const uint64_t ui = c2;
            used = (uint32_t)ui;
            inst = (uint32_t)(ui >> 32);
            const uint32_t open = (uint32_t)(c1 >> 56);
            if (open) {
                // the open window and the candidates at its lowest open cell,
                // rebuilt (as the level before computed them)
                cov = c1 & ((1ull << 56) - 1);
                E = (row_window(inst) & ~(cov << 8) & ~0xffull) | open;
                const uint32_t c = (uint32_t)__builtin_ctz(open);
                av = avail_of(s_xa, used);
                ok = fitmask<FR>(s_fit, a.fit_rows, c, E) & av;
            }
        }
        if (FROM_REC) {
            uint64_t key2 = cov | ((uint64_t)used << a.occ_bits);
            if (a.inst_shift < 64) key2 |= (uint64_t)inst << a.inst_shift;
            if (!emit_children(pass, key2, cnt)) {
                alive = false;
                break;
            }
        }
        // the loaded rows go to LDS; their pairs are listed pass by pass
        __builtin_amdgcn_wave_barrier();
        o_E[lane] = E;
        o_cov[lane] = cov;
        o_cnt[lane] = cnt;
        o_ui[lane] = (uint64_t)used | ((uint64_t)inst << 32);
        o_av[lane] = ok ? av : 0ull;
        const uint32_t n = (uint32_t)__builtin_popcountll(ok);
        const uint32_t incl = wave_incl_scan(n);
        const uint32_t excl = incl - n;
        bool pending = n != 0;
        while (alive) {
            const uint64_t pendm = __ballot(pending);
            if (!pendm) break;
            const uint32_t first = (uint32_t)__builtin_ctzll(pendm);
            const uint32_t off = __shfl(excl, (int)first);
            const bool fits = pending && incl - off <= PCAP;
            const uint64_t fm = __ballot(fits);
            const uint32_t lastf = 63u - (uint32_t)__builtin_clzll(fm);
            const uint32_t total = __shfl(incl, (int)lastf) - off;
            __builtin_amdgcn_wave_barrier();
            if (fits) {
                uint32_t t = excl - off;
                for (uint64_t m = ok; m; m &= m - 1)
                    plist[t++] = (uint16_t)((lane << 6) | (uint32_t)__builtin_ctzll(m));
            }
            __builtin_amdgcn_wave_barrier();
            pending = pending && !fits;
            for (uint32_t r0 = 0; r0 < total && alive; r0 += 64) {
                const uint32_t p = r0 + lane;
                bool emitC = false, emitS = false;
                uint64_t E2 = 0, cov2 = 0, ok2 = 0, pc = 0;
                uint32_t used2 = 0, pi = 0;
                if (p < total) {
                    const uint32_t e = plist[p];
                    const uint32_t o = e >> 6, k = e & 63u;
                    const uint64_t pE = o_E[o];
                    const uint64_t pui = o_ui[o];
                    const uint32_t c = (uint32_t)__builtin_ctzll(pE & 0xffu);
                    const PlaceInfo pk = s_pc[c * MAX_CAND + k];
                    const uint64_t pl = pk.shape;
                    if (!(pl & ~pE)) {
                        const uint32_t ci = (uint32_t)pk.info;
                        E2 = pE & ~pl;
                        used2 = (uint32_t)pui + (1u << (ci & 0xffu));
                        pi = (uint32_t)(pui >> 32);
                        cov2 = o_cov[o] | (pl >> 8);
                        pc = o_cnt[o];
                        if (!(E2 & 0xffu)) {
                            emitC = true;
                        } else {
                            uint64_t av2 = o_av[o];
                            if (((used2 >> (ci & 0xffu)) & ((ci >> 8) & 0xffu)) == (ci >> 16))
                                av2 &= ~s_rng[k];
                            const uint32_t c2 = (uint32_t)__builtin_ctzll(E2 & 0xffu);
                            ok2 = fitmask<FR>(s_fit, a.fit_rows, c2, E2) & av2;
                            emitS = ok2 != 0;
                        }
                    }
                }
                {
                    uint64_t key2 = cov2 | ((uint64_t)used2 << a.occ_bits);
                    if (a.inst_shift < 64) key2 |= (uint64_t)pi << a.inst_shift;
                    if (!emit_children(emitC, key2, pc)) {
                        alive = false;
                        break;
                    }
                }
                const uint64_t sm = __ballot(emitS);
                if (sm) {
                    bool okw = true;
                    const uint64_t sl = place(ws, sm, ctr_s, chunk, a.out_cap, a.c32, okw);
                    if (!okw) {
                        alive = false;
                        break;
                    }
                    if (emitS) {
                        oCnt[sl] = pc;
                        oCw[sl] = cov2 | ((E2 & 0xffull) << 56);
                        oUi[sl] = (uint64_t)used2 | ((uint64_t)pi << 32);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // (the children still staged)
    if (alive && stage && n_s && !flush_staged(n_s)) alive = false;
    if (!alive) return;  // the step is rerun
    // unused slots of the wave's last chunks: no candidates / empty keys
    for (uint64_t j = ws.pos + lane; j < ws.end; j += 64) oCw[j] = 0;
    for (uint64_t j = wc.pos + lane; j < wc.end; j += 64) a.ch_key[j] = EMPTY_KEY;
    pruned = wave_sum_u64(pruned);
    if (lane == 0 && pruned) atomicAdd(a.c64 + C64_LPRUNED, (unsigned long long)pruned);
}

// Board windows of the children's row, per instance, staged in LDS when the
// batch has at most this many instances.
constexpr uint32_t DEDUP_LDS_INST = 1024;
constexpr uint32_t DEDUP_UNROLL = 4;  // children per thread per pass (their probes overlap)

// Merges one child into the table, its first probe already loaded (cur):
// returns 1 new, 0 merged, -1 no slot within the probe limit.  A slot
// holding the key only takes the count (one atomic add, no CAS); an empty
// slot is claimed by CAS; keys only ever change EMPTY -> key, so a load that
// saw another key lets the probe move on.
__device__ __forceinline__ int dedup_insert(const DedupArgs &a, uint64_t key, uint64_t cnt,
                                            uint64_t h, uint64_t cur) {
    for (uint32_t p = 0;;) {
        if (cur == EMPTY_KEY) {
            cur = atomicCAS(a.tab + 2 * h, (unsigned long long)EMPTY_KEY, (unsigned long long)key);
            if (cur == EMPTY_KEY) {
                atomicAdd(a.tab + 2 * h + 1, (unsigned long long)cnt);
                return 1;
            }
        }
        if (cur == key) {
            atomicAdd(a.tab + 2 * h + 1, (unsigned long long)cnt);
            return 0;
        }
        if (++p >= a.max_probe) return -1;
        h = next_slot(h, a.tab_size);
        cur = __hip_atomic_load(a.tab + 2 * h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool PRUNE, bool LDSW>
__global__ __launch_bounds__(256) void k_dedup(DedupArgs a) {
    __shared__ uint64_t s_win[LDSW ? DEDUP_LDS_INST : 1];
    if (a.ctl) {
        if (a.ctl->fail) return;
        a.row = a.ctl->row;
        a.tab_size = min((uint64_t)a.ctl->tab_m, a.tab_size);
    }
    if (LDSW) {
        for (uint32_t i = threadIdx.x; i < a.n_inst; i += 256) {
            const uint32_t rr = a.row + 1 - a.start[i];
            s_win[i] = rr < a.bw_stride ? a.bw[(size_t)i * a.bw_stride + rr] : 0ull;
        }
        __syncthreads();
    }
    uint64_t n = ctr_load(a.c64 + C64_CHW);
    if (n > a.ch_cap) n = a.ch_cap;
    // a chunk whose level kernels overflowed is rerun by the host: nothing
    // of it enters the table; a complete chunk counts itself once
    if (__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        n = 0;
    } else if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(a.c32 + C32_LVDONE, 1u);
        // the children the level kernels pruned count now that the chunk is
        // complete (once per row step when a streamed step reruns the level
        // kernels for each partition of the hash)
        const uint64_t lp = ctr_load(a.c64 + C64_LPRUNED);
        if (lp && (!a.part_mask || a.part_id == 0))
            atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)lp);
    }
    uint64_t kept = 0, fresh = 0, pruned = 0;
    bool failed = false;
    const uint64_t G = (uint64_t)gridDim.x * 256;
    for (uint64_t i0 = blockIdx.x * 256ull + threadIdx.x; i0 < n; i0 += DEDUP_UNROLL * G) {
        uint64_t key[DEDUP_UNROLL], cnt[DEDUP_UNROLL], h[DEDUP_UNROLL], cur[DEDUP_UNROLL];
        bool live[DEDUP_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < DEDUP_UNROLL; u++) {
            const uint64_t i = i0 + u * G;
            key[u] = i < n ? a.ch_key[i] : EMPTY_KEY;
            cnt[u] = i < n ? a.ch_cnt[i] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < DEDUP_UNROLL; u++) {
            live[u] = key[u] != EMPTY_KEY;
            // (children of other partitions are left to their own pass)
            if (a.part_mask && live[u] && (mix64(key[u]) & a.part_mask) != a.part_id)
                live[u] = false;
            if (PRUNE && live[u]) {
                const uint32_t inst = a.inst_shift < 64 ? (uint32_t)(key[u] >> a.inst_shift) : 0u;
                uint64_t w;
                if (LDSW) {
                    w = s_win[inst];
                } else {
                    const uint32_t rr = a.row + 1 - a.start[inst];
                    w = a.bw[(size_t)inst * a.bw_stride + rr];
                }
                const uint64_t E = w & ~(key[u] & a.occ_mask);
                const bool ok = a.closed_rows <= 4 ? viable32(E, a.closed_rows, AreaMask{a.area_ok})
                                                   : viable(E, a.seed_mask, AreaMask{a.area_ok});
                if (!ok) {
                    pruned++;
                    live[u] = false;
                }
            }
            if (live[u]) {
                h[u] = home_slot(key[u], a.tab_size);
                cur[u] = __hip_atomic_load(a.tab + 2 * h[u], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < DEDUP_UNROLL; u++) {
            if (!live[u]) continue;
            kept++;
            int r = dedup_insert(a, key[u], cnt[u], h[u], cur[u]);
            if (r < 0 && a.spill_cap) {
                // no slot within the probe limit: park it for the host's table growth
                const uint32_t s = atomicAdd(a.c32 + C32_SPILL, 1u);
                if (s < a.spill_cap) {
                    a.spill[2 * s] = key[u];
                    a.spill[2 * s + 1] = cnt[u];
                    r = 0;
                }
            }
            if (r < 0) failed = true;
            fresh += r > 0;
        }
    }
    if (failed) atomicOr(a.c32 + C32_OVERFLOW, 1u);
    kept = wave_sum_u64(kept);
    fresh = wave_sum_u64(fresh);
    pruned = wave_sum_u64(pruned);
    if (lane_id() == 0) {
        if (kept) atomicAdd(a.c64 + C64_CHILDREN, (unsigned long long)kept);
        if (fresh) atomicAdd(a.c64 + C64_NEW, (unsigned long long)fresh);
        if (pruned) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)pruned);
    }
}

template <bool FROM_REC>
static void launch_lv(const LevelArgs &a, dim3 g, size_t lds, hipStream_t st) {
    switch (a.fit_rows) {
        case 1: hipLaunchKernelGGL((k_level<FROM_REC, 1>), g, dim3(LT), lds, st, a); break;
        case 2: hipLaunchKernelGGL((k_level<FROM_REC, 2>), g, dim3(LT), lds, st, a); break;
        case 3: hipLaunchKernelGGL((k_level<FROM_REC, 3>), g, dim3(LT), lds, st, a); break;
        default: hipLaunchKernelGGL((k_level<FROM_REC, 0>), g, dim3(LT), lds, st, a); break;
    }
}

}  // namespace

size_t level_lds_bytes(uint32_t fit_rows, uint32_t n_win, uint32_t n_win0, bool stage) {
    return sizeof(uint64_t) * (8u * fit_rows * 256u + 2u * 8u * MAX_CAND + 4u * 256u + MAX_CAND) +
           (sizeof(uint64_t) * OWN * 64 + sizeof(uint16_t) * PCAP) * (LT / 64) +
           sizeof(uint64_t) * (n_win + n_win0 + (n_win0 + 7) / 8) +
           (stage ? sizeof(uint64_t) * 128 * (LT / 64) : 0);
}

void launch_level(const LevelArgs &a, int n_cu, hipStream_t st) {
    const size_t lds =
        level_lds_bytes(a.fit_rows, a.prune ? a.n_win : 0, a.n_win0, a.prune && a.stage);
    // one workgroup per CU (LDS-bound); level 0 knows its record count
    uint64_t blocks = (uint64_t)n_cu;
    if (a.level == 0) blocks = std::min<uint64_t>(blocks, ((uint64_t)a.n_rec + LT - 1) / LT);
    if (!blocks) blocks = 1;
    if (a.level == 0) launch_lv<true>(a, dim3((uint32_t)blocks), lds, st);
    else launch_lv<false>(a, dim3((uint32_t)blocks), lds, st);
}

void launch_dedup(const DedupArgs &a, bool prune, int n_cu, hipStream_t st) {
    const dim3 g((uint32_t)n_cu * 8);
    if (!prune) hipLaunchKernelGGL((k_dedup<false, false>), g, dim3(256), 0, st, a);
    else if (a.n_inst <= DEDUP_LDS_INST) hipLaunchKernelGGL((k_dedup<true, true>), g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_dedup<true, false>), g, dim3(256), 0, st, a);
}

}  // namespace d4
