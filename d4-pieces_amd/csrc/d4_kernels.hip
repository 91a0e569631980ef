// d4_kernels.hip -- gfx950 kernels of the row-by-row polyomino tiling search.
//
// One row step turns the frontier of row r (states: occupancy of the rows
// below r by pieces already placed + per-piece use counters, with a u64
// tiling count) into the frontier of row r+1:
//
//   k_extend  each lane takes one parent record and enumerates, depth first,
//             every way to fill the remaining open cells of row r with
//             pieces anchored at the lowest open cell (the reference CPU
//             searcher's order, src/Piece.cpp:82-98).  Candidates at an
//             anchor come from byte-sliced fit tables staged in LDS:
//             fit[c][j][byte] = mask of (piece, transform) pairs anchored at
//             column c whose row-j cells lie inside `byte`, so one AND per
//             window row (the first fit_rows rows) leaves the placements
//             that fit there; a full-shape test on the chosen one covers the
//             rows below.  Every
//             completed row is a child: it is checked for enclosed regions
//             that no piece set can fill (connectivity pruning) and merged
//             into an open-addressing hash table keyed by the packed state,
//             summing counts (dedup).  In enum mode each child also writes a
//             back-pointer edge, slots claimed by wave ballot + prefix count
//             with one atomic per wave.
//   k_compact scans the hash table and writes the unique states densely
//             (wave ballot + mbcnt, one atomic per wave).
//
// This replaces the reference's tiled_row_search / impl<H> / host Sorter
// pipeline (src/kernel.cuh:12-160, src/sorter.cpp:156-214): there the L x R
// half-row cross product is tested per record, every child takes one
// atomicAdd, and dedup runs on the host over PCIe.
#include "d4_internal.h"
#include "d4_device.h"

#include <algorithm>

namespace d4 {

namespace {

using namespace dev;

constexpr uint32_t STAGE = 128;  // staged children per wave
constexpr int CMP_THREADS = 256;

// COUNT_ONLY runs the same search but only counts the surviving children
// (used to size the dedup table exactly when the frontier is small).
template <bool ENUM, bool PRUNE, bool COUNT_ONLY, int THREADS, int FR>
__global__ __launch_bounds__(THREADS) void k_extend(ExtendArgs a) {
    constexpr uint32_t SW = ENUM ? 5 : 3;  // staged u64 words per child
    extern __shared__ uint64_t lds[];
    const uint32_t nfit = 8u * a.fit_rows * 256u;
    uint64_t *s_fit = lds;
    const PlaceInfo *s_pc = reinterpret_cast<const PlaceInfo *>(s_fit + nfit);
    uint64_t *s_xa = s_fit + nfit + 2 * 8 * MAX_CAND;
    uint64_t *s_rng = s_xa + 4 * 256;  // per candidate: its piece's candidate range
    uint64_t *s_stage = s_rng + MAX_CAND;
    for (uint32_t i = threadIdx.x; i < nfit; i += THREADS) s_fit[i] = a.fit[i];
    for (uint32_t i = threadIdx.x; i < 2 * 8 * MAX_CAND; i += THREADS)
        s_fit[nfit + i] = a.placeinfo[i];
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += THREADS) s_xa[i] = a.xavail[i];
    for (uint32_t i = threadIdx.x; i < MAX_CAND; i += THREADS)
        s_rng[i] = piece_range(a.placeinfo[2 * i + 1]);
    __syncthreads();

    const uint32_t lane = lane_id();
    const uint32_t rows = a.fit_rows;
    uint64_t n_children = 0, n_new = 0, n_pruned = 0;

    // Work distribution.  A work item is a record, or with split levels
    // fix = 1 / 2 a (record, first placement[, second placement]) triple --
    // 64 / 4096 items per record -- so tiny frontiers (the first rows) still
    // fill the GPU.  A wave claims 64 consecutive items with one atomic and
    // loads the records they cover with one coalesced access (lane j holds
    // record j of the chunk with its board windows); lanes whose search is
    // over take the next item of the chunk by shuffle, so a record with a
    // large subtree does not idle the rest of its wave and no lane waits on
    // memory per record.
    const bool don_mode = a.don_in != nullptr;
    const uint32_t fix = don_mode ? 0u : a.split;
    const uint32_t per_rec = 1u << (6 * fix);
    // a follow-up launch queued without a host round trip reads its item
    // count from the counters of the launch before it
    uint32_t n_don = a.n_don;
    if (don_mode && n_don == ~0u) {
        const uint32_t dn = ctr_load(a.c32 + a.don_read), de = ctr_load(a.c32 + a.don_read + 1);
        n_don = dn <= a.don_cap ? dn : (de ? de - 1 : a.don_cap);
    }
    const uint64_t n_items = don_mode ? (uint64_t)n_don : (uint64_t)a.n_in * per_rec;
    bool exhausted = false;
    uint32_t buf_items = 0, buf_pos = 0, buf_rec0 = 0;
    uint64_t buf_item0 = 0;
    uint64_t bkey = 0, bcnt = 0, bBW = 0, bBWn = 0;
    uint32_t bterm = 0;
    uint64_t bit0 = 0, bit1 = 0;  // donated item words (don_mode)

    bool searching = false, pass = false;
    uint32_t rec = 0, inst = 0, used = 0, d = 0, c = 0, anc = 0;
    // the lane's work item owns levels >= base_d of the DFS; sib: whether
    // siblings at level base_d belong to it; steps: DFS steps on the item
    uint32_t base_d = 0, steps = 0;
    bool sib = true, no_don = false;
    uint64_t cnt = 0, BW = 0, BWn = 0, E = 0, avail = 0, desc = 0, ok = 0;

    // Children are staged per wave in LDS (key, count, open window; in enum
    // mode also the row descriptor and the parent) and flushed 64 at a time:
    // each lane prunes and inserts one, and enum mode writes the kept
    // children's edges with one slot allocation per flush.
    uint64_t *st_key = s_stage + (threadIdx.x >> 6) * (SW * STAGE);  // per-wave stage
    uint64_t *st_cnt = st_key + STAGE;
    uint64_t *st_E = st_cnt + STAGE;
    uint64_t *st_desc = st_E + STAGE;     // enum mode: row descriptor
    uint64_t *st_rec = st_desc + STAGE;   // enum mode: parent record
    uint32_t sc = 0;  // staged entries (wave-uniform)
    // Count mode pipelines the table insert: a flush issues the first probe
    // (the CAS) of each kept child and leaves it in flight; the next flush
    // settles it (merge or claim, further probes, spill), so the round trip
    // overlaps the DFS steps in between instead of stalling the wave.
    constexpr bool PIPE = !ENUM && !COUNT_ONLY;
    bool p_on = false;
    uint64_t p_key = 0, p_cnt = 0, p_h = 0, p_cur = 0;
    // Settles this lane's pending insert: returns tally bit 2 (new) and sets
    // `failed` when no slot was found within the probe limit and the spill
    // list is full.
    auto settle = [&](bool &failed) __attribute__((always_inline)) -> uint32_t {
        if (!p_on) return 0u;
        p_on = false;
        int r = -1;
        uint64_t h = p_h, cur = p_cur;
        for (uint32_t pr = 1;; pr++) {
            if (cur == EMPTY_KEY || cur == p_key) {
                atomicAdd(a.tab_cnt + 2 * h, (unsigned long long)p_cnt);
                r = cur == EMPTY_KEY ? 1 : 0;
                break;
            }
            if (pr >= a.max_probe) break;
            h = next_slot(h, a.tab_size);
            cur = atomicCAS(a.tab_key + 2 * h, (unsigned long long)EMPTY_KEY,
                            (unsigned long long)p_key);
        }
        if (r < 0 && a.spill_cap) {
            const uint32_t sp = atomicAdd(a.c32 + C32_SPILL, 1u);
            if (sp < a.spill_cap) {
                a.spill[2 * sp] = p_key;
                a.spill[2 * sp + 1] = p_cnt;
                r = 0;
            }
        }
        if (r < 0) failed = true;
        return r > 0 ? 4u : 0u;
    };
    // Returns this lane's tallies: bit 0 child kept, bit 1 pruned, bit 2 new.
    auto flush = [&](uint32_t n) __attribute__((always_inline)) -> uint32_t {
        __builtin_amdgcn_wave_barrier();
        const bool has = lane < n;
        const uint32_t sl = sc - n + lane;
        const uint64_t k2 = has ? st_key[sl] : 0ull;
        const uint64_t c2 = has ? st_cnt[sl] : 0ull;
        const uint64_t e2 = has ? st_E[sl] : 0ull;
        const uint64_t d2 = ENUM && has ? st_desc[sl] : 0ull;
        const uint32_t r2 = ENUM && has ? (uint32_t)st_rec[sl] : 0u;
        __builtin_amdgcn_wave_barrier();
        sc -= n;
        bool keep = has;
        uint32_t t = 0;
        if (PRUNE && has) {
            keep = viable(e2, a.seed_mask, a.gcd, a.min_size);
            t |= keep ? 0u : 2u;
        }
        bool failed = false;
        uint64_t slot = 0;
        if (PIPE) t |= settle(failed);
        if (keep) {
            t |= 1u;
            if (PIPE) {
                p_h = home_slot(k2, a.tab_size);
                p_cur = atomicCAS(a.tab_key + 2 * p_h, (unsigned long long)EMPTY_KEY,
                                  (unsigned long long)k2);
                p_key = k2;
                p_cnt = c2;
                p_on = true;
            } else if (!COUNT_ONLY) {
                int r = table_insert(a, k2, c2, slot);
                if (r < 0 && !ENUM && a.spill_cap) {
                    // no slot within the probe limit: park the child for the
                    // host's table growth (counted as new there)
                    const uint32_t sp = atomicAdd(a.c32 + C32_SPILL, 1u);
                    if (sp < a.spill_cap) {
                        a.spill[2 * sp] = k2;
                        a.spill[2 * sp + 1] = c2;
                        r = 0;
                    }
                }
                if (r < 0) failed = true;
                else t |= r ? 4u : 0u;
            }
        }
        if (ENUM && !COUNT_ONLY) {
            const bool edge = keep && !failed;
            const uint32_t e = wave_alloc(a.c32 + C32_EDGES, edge);
            if (edge) {
                if (e < a.e_cap) {
                    a.e_parent[e] = a.rec_base + r2;
                    a.e_desc[e] = d2;
                    a.e_slot[e] = slot;
                } else {
                    atomicOr(a.c32 + C32_EDGE_OVF, 1u);
                }
            }
        }
        if (__any(failed)) {
            // table too small: flag it and stop; the host reruns the step
            if (failed) atomicOr(a.c32 + C32_OVERFLOW, 1u);
            searching = false;
            pass = false;
            exhausted = true;
            buf_pos = buf_items;
        }
        return t;
    };
    auto tally = [&](uint32_t t) __attribute__((always_inline)) {
        n_children += t & 1u;
        n_pruned += (t >> 1) & 1u;
        n_new += (t >> 2) & 1u;
    };
    // Hand the rest of this lane's search tree to the next launch: the
    // untried candidates at the current level and, at each level below it
    // that the item owns, the candidates after the one in progress.  The
    // items cover exactly the work not yet done.  Returns false (and the
    // lane goes on) when there is nothing to hand over or no room.
    auto donate = [&]() __attribute__((always_inline)) -> bool {
        // the current level's untried candidates go out one item each (the
        // widest fan-out), the levels below as "candidates after k" items
        uint32_t n = (uint32_t)__builtin_popcountll(ok);
        for (uint32_t l = base_d; l < d; l++) {
            const uint32_t kl = (uint32_t)((desc >> (8 * l)) & 0xffu) - 1u;
            n += (l > base_d || sib) && kl + 1 < (uint32_t)MAX_CAND;
        }
        if (!n) return false;
        // once the buffer is full nobody adds to the counter again (so it
        // cannot wrap), and a lane that found no room keeps its item
        if (__hip_atomic_load(a.c32 + a.don_write, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
            a.don_cap) {
            no_don = true;
            atomicAdd(a.c32 + C32_DONFULL, 1u);
            return false;
        }
        const uint32_t b = atomicAdd(a.c32 + a.don_write, n);
        if ((uint64_t)b + n > (uint64_t)a.don_cap) {
            no_don = true;
            if (b < a.don_cap) a.c32[a.don_write + 1] = b + 1;  // the one range crossing the end
            atomicAdd(a.c32 + C32_DONFULL, 1u);
            return false;
        }
        atomicAdd(a.c64 + C64_DONATED, (unsigned long long)n);
        const uint64_t r = (uint64_t)(a.rec_base + rec);
        uint32_t o = b;
        auto put = [&](uint32_t lvl, uint32_t first, bool single) __attribute__((always_inline)) {
            const uint64_t pm = lvl >= 8 ? ~0ull : (1ull << (8 * lvl)) - 1ull;
            a.don_out[2 * o] = r | ((uint64_t)first << 32) | ((uint64_t)lvl << 40) |
                               ((uint64_t)single << 48);
            a.don_out[2 * o + 1] = desc & pm;
            o++;
        };
        for (uint64_t m = ok; m; m &= m - 1) put(d, (uint32_t)__builtin_ctzll(m), true);
        for (uint32_t l = base_d; l < d; l++) {
            const uint32_t kl = (uint32_t)((desc >> (8 * l)) & 0xffu) - 1u;
            if ((l > base_d || sib) && kl + 1 < (uint32_t)MAX_CAND) put(l, kl + 1, false);
        }
        return true;
    };

    for (;;) {
        // refill idle lanes once at least a.refill of them are idle (or the
        // whole wave is): each refill round costs the wave a full handout
        uint64_t im = __ballot(!(searching || pass));
        if ((uint32_t)__builtin_popcountll(im) < a.refill && ~im) im = 0;
        while (im) {
            if (buf_pos >= buf_items) {
                if (exhausted) break;
                uint64_t base = 0;
                if (lane == 0) {
                    if (__hip_atomic_load(a.c32 + C32_OVERFLOW, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT))
                        base = n_items;  // table overflow: the step is rerun
                    else
                        base = atomicAdd(a.c64 + C64_WORK, 64ull);
                }
                base = shfl64(base, 0);
                if (base >= n_items) {
                    exhausted = true;
                    break;
                }
                buf_item0 = base;
                buf_items = (uint32_t)min((uint64_t)64, n_items - base);
                buf_pos = 0;
                buf_rec0 = (uint32_t)(base >> (6 * fix));
                const uint32_t nrec = fix ? 1u : buf_items;
                if (lane < nrec) {
                    uint32_t r0 = buf_rec0 + lane;
                    bool bad = false;
                    if (don_mode) {
                        bit0 = a.don_in[2 * (buf_item0 + lane)];
                        bit1 = a.don_in[2 * (buf_item0 + lane) + 1];
                        r0 = (uint32_t)bit0;
                        bad = r0 >= a.n_in;  // never expected: flagged, skipped
                        if (bad) {
                            atomicOr(a.c32 + C32_BAD_ITEM, 1u);
                            r0 = 0;
                        }
                    }
                    bkey = a.in_key[r0];
                    bcnt = a.in_cnt[r0];
                    // the instance's own row: instances join at their start step
                    uint32_t bi = a.inst_shift < 64 ? (uint32_t)(bkey >> a.inst_shift) : 0u;
                    const uint32_t rr = a.row - a.start[bi];
                    bterm = bad || rr >= a.hgt[bi];
                    const uint64_t *w = a.bw + (size_t)bi * a.bw_stride + rr;
                    bBW = bterm ? 0 : w[0];
                    bBWn = bterm ? 0 : w[1];
                }
            }
            const bool idle = (im >> lane) & 1u;
            const uint32_t q = buf_pos + mbcnt(im);
            const uint32_t take = min((uint32_t)__builtin_popcountll(im), buf_items - buf_pos);
            buf_pos += take;
            // every lane joins the shuffles; idle lanes with q in range keep the result
            const uint32_t src = fix ? 0u : min(q, 63u);
            const uint64_t key = shfl64(bkey, src);
            const uint64_t kc = shfl64(bcnt, src);
            const uint64_t kBW = shfl64(bBW, src);
            const uint64_t kBWn = shfl64(bBWn, src);
            const uint32_t kterm = __shfl(bterm, (int)src);
            uint64_t it0 = 0, it1 = 0;
            if (don_mode) {
                it0 = shfl64(bit0, src);
                it1 = shfl64(bit1, src);
            }
            if (idle && q < buf_items) {
                const uint64_t item = buf_item0 + q;
                const uint32_t sub = (uint32_t)(item & (per_rec - 1));  // k0 << 6 | k1
                const bool lead_item = sub == 0 && !don_mode;
                rec = don_mode ? (uint32_t)it0 : buf_rec0 + (fix ? 0u : q);
                steps = 0;
                no_don = false;
                base_d = 0;
                sib = fix == 0;
                cnt = kc;
                inst = a.inst_shift < 64 ? (uint32_t)(key >> a.inst_shift) : 0u;
                used = (uint32_t)(key >> a.occ_bits) & a.used_mask;
                if (kterm) {
                    // terminal state of a finished instance: all cells covered
                    if (lead_item && !COUNT_ONLY)
                        atomicAdd(a.result + inst, (unsigned long long)cnt);
                } else {
                    BW = kBW;
                    BWn = kBWn;
                    E = BW & ~(key & a.occ_mask);
                    d = 0;
                    anc = 0;
                    desc = 0;
                    if (E & 0xffu) {
                        avail = avail_of(s_xa, used);
                        c = (uint32_t)__builtin_ctzll(E & 0xffu);
                        ok = fitmask<FR>(s_fit, rows, c, E) & avail;
                        if (fix == 1) {
                            ok &= 1ull << sub;
                        } else if (fix == 2) {
                            // commit the first placement k0, keep only k1 next
                            const uint32_t k0 = sub >> 6, k1 = sub & 63u;
                            ok &= 1ull << k0;
                            const PlaceInfo pi0 = s_pc[c * MAX_CAND + k0];
                            const uint64_t pl = pi0.shape;
                            if (ok && (pl & ~E)) ok = 0;
                            if (ok) {
                                const uint32_t ci = (uint32_t)pi0.info;
                                const uint64_t E1 = E & ~pl;
                                if (!(E1 & 0xffu)) {
                                    // k0 completes the row: the k1 == 0 item
                                    // emits that child on its own
                                    ok = k1 == 0 ? ok : 0;
                                } else {
                                    E = E1;
                                    used += 1u << (ci & 0xffu);
                                    if (((used >> (ci & 0xffu)) & ((ci >> 8) & 0xffu)) == (ci >> 16))
                                        avail &= ~s_rng[k0];
                                    desc = (uint64_t)(k0 + 1);
                                    anc = c;
                                    d = 1;
                                    base_d = 1;
                                    c = (uint32_t)__builtin_ctzll(E & 0xffu);
                                    ok = fitmask<FR>(s_fit, rows, c, E) & avail & (1ull << k1);
                                }
                            }
                        } else if (don_mode) {
                            // replay the item's placements, then take the
                            // candidates from `first` on at its level
                            const uint32_t lvl = (uint32_t)(it0 >> 40) & 0xffu;
                            const uint32_t first = (uint32_t)(it0 >> 32) & 0xffu;
                            const bool single = (it0 >> 48) & 1u;
                            for (uint32_t j = 0; j < lvl; j++) {
                                const uint32_t kj = ((uint32_t)((it1 >> (8 * j)) & 0xffu) - 1u) & 63u;
                                const PlaceInfo pj = s_pc[c * MAX_CAND + kj];
                                const uint32_t cj = (uint32_t)pj.info;
                                E &= ~pj.shape;
                                used += 1u << (cj & 0xffu);
                                if (((used >> (cj & 0xffu)) & ((cj >> 8) & 0xffu)) == (cj >> 16))
                                    avail &= ~s_rng[kj];
                                anc |= c << (3 * j);
                                c = (uint32_t)__builtin_ctzll(E & 0xffu);
                            }
                            desc = it1;
                            d = lvl;
                            base_d = lvl;
                            sib = !single;
                            ok = fitmask<FR>(s_fit, rows, c, E) & avail &
                                 (single ? 1ull << first : ~0ull << first);
                        }
                        searching = ok != 0;
                    } else {
                        pass = lead_item;  // row already covered from above
                    }
                }
            }
            im = __ballot(!(searching || pass));
        }
        if (!__any(searching || pass)) break;

        // one DFS step over the completions of row r.  Stack: desc holds k+1
        // per level (one byte each), anc the anchor column (3 bits each).
        bool emit = false;
        uint64_t E2 = 0, desc2 = 0;
        uint32_t used2 = 0;
        if (pass) {
            emit = true;
            E2 = E;
            used2 = used;
            pass = false;
        } else if (searching) {
            if (a.budget && !no_don && ++steps > a.budget && donate()) {
                searching = false;
            } else if (ok) {
                uint32_t k = (uint32_t)__builtin_ctzll(ok);
                ok &= ok - 1;
                const PlaceInfo pk = s_pc[c * MAX_CAND + k];
                const uint64_t pl = pk.shape;
                const uint32_t ci = (uint32_t)pk.info;
                E2 = E & ~pl;
                uint32_t inc = 1u << (ci & 0xffu);
                if (pl & ~E) {
                    // fails in a window row the fit table does not cover
                } else if (!(E2 & 0xffu)) {
                    emit = true;
                    used2 = used + inc;
                    desc2 = desc | ((uint64_t)(k + 1) << (8 * d));
                } else {
                    E = E2;
                    used += inc;
                    if (((used >> (ci & 0xffu)) & ((ci >> 8) & 0xffu)) == (ci >> 16))
                        avail &= ~s_rng[k];
                    desc |= (uint64_t)(k + 1) << (8 * d);
                    anc |= c << (3 * d);
                    d++;
                    c = (uint32_t)__builtin_ctzll(E & 0xffu);
                    ok = fitmask<FR>(s_fit, rows, c, E) & avail;
                }
            } else if (d <= base_d) {
                // the levels below base_d are fixed by the work item
                searching = false;
            } else {
                d--;
                uint32_t k = (uint32_t)((desc >> (8 * d)) & 0xffu) - 1u;
                c = (anc >> (3 * d)) & 7u;
                desc &= ~(0xffull << (8 * d));
                anc &= ~(7u << (3 * d));
                const PlaceInfo pk = s_pc[c * MAX_CAND + k];
                E |= pk.shape;
                used -= 1u << (pk.info & 0xffu);
                avail |= s_rng[k];
                ok = (d > base_d || sib) ? fitmask<FR>(s_fit, rows, c, E) & avail & ((~0ull << k) << 1)
                                         : 0;
            }
        }
        // queue the child in the wave's LDS stage: pruning, the table insert
        // and the edge write run in flush(), 64 children at a time with
        // every lane active, instead of stalling this divergent loop on each
        // CAS
        const uint64_t em = __ballot(emit);
        if (em) {
            if (emit) {
                const uint32_t sl = sc + mbcnt(em);
                uint64_t key2 = ((BW & ~E2) >> 8) | ((uint64_t)used2 << a.occ_bits);
                if (a.inst_shift < 64) key2 |= (uint64_t)inst << a.inst_shift;
                st_key[sl] = key2;
                st_cnt[sl] = cnt;
                st_E[sl] = (E2 >> 8) | (BWn & 0xff00000000000000ull);
                if (ENUM) {
                    st_desc[sl] = desc2;
                    st_rec[sl] = rec;
                }
            }
            sc += (uint32_t)__builtin_popcountll(em);
            if (sc >= 64) tally(flush(64));
        }
    }
    while (sc) tally(flush(min(sc, 64u)));
    if (PIPE) {
        bool failed = false;
        tally(settle(failed));
        if (failed) atomicOr(a.c32 + C32_OVERFLOW, 1u);
    }

    n_children = wave_sum_u64(n_children);
    n_new = wave_sum_u64(n_new);
    n_pruned = wave_sum_u64(n_pruned);
    if (lane == 0) {
        if (n_children) atomicAdd(a.c64 + C64_CHILDREN, (unsigned long long)n_children);
        if (n_new) atomicAdd(a.c64 + C64_NEW, (unsigned long long)n_new);
        if (n_pruned) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)n_pruned);
    }
}

__global__ __launch_bounds__(256) void k_fill_table(unsigned long long *key,
                                                    unsigned long long *cnt,
                                                    uint64_t n, const RowCtl *ctl) {
    if (ctl) {
        if (ctl->fail) return;
        n = min(n, (uint64_t)ctl->tab_m);
    }
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        key[2 * i] = EMPTY_KEY;
        cnt[2 * i] = 0;
    }
}

// 4-word slots {k0, k1, count, unused} of the multi-library batch, written
// whole (two 16-byte stores per slot), so every cache line is written in full.
__global__ __launch_bounds__(256) void k_fill_table4(unsigned long long *tab, uint64_t n) {
    ulonglong2 *t2 = reinterpret_cast<ulonglong2 *>(tab);
    const ulonglong2 keys = make_ulonglong2(EMPTY_KEY, EMPTY_KEY), zero = make_ulonglong2(0, 0);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        t2[2 * i] = keys;
        t2[2 * i + 1] = zero;
    }
}

__global__ void k_reset_counters(uint32_t *c32, unsigned long long *c64, const RowCtl *ctl) {
    if (ctl && ctl->fail) return;
    if (threadIdx.x < C32_N) c32[threadIdx.x] = 0;
    if (threadIdx.x < C64_N) c64[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_copy_u64(unsigned long long *dst,
                                                  const unsigned long long *src, uint64_t n,
                                                  const RowCtl *ctl) {
    if (ctl && ctl->fail) return;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

// The last kernel of a graph-replayed row step (one lane): a step whose level
// buffers, table or spill list overflowed, or whose compaction count differs
// from its new states, fails and stays put (the frontier it read is intact
// and the steps after it do nothing); otherwise it is logged and the next
// step's frontier, row and table size are set -- the table sized as the host
// sizes a count-mode step from the unique states per parent, capped at the
// allocation.
__global__ void k_row_close(RowCtl *ctl, const uint32_t *c32, const unsigned long long *c64) {
    if (threadIdx.x != 0 || ctl->fail) return;
    const uint64_t n = ctl->n_front, u = ctr_load(c64 + C64_NEW);
    uint32_t f = (ctr_load(c32 + C32_XOVF) ? 1u : 0u) | (ctr_load(c32 + C32_OVERFLOW) ? 2u : 0u) |
                 (ctr_load(c32 + C32_SPILL) ? 4u : 0u) | (ctr_load(c32 + C32_POVF) ? 16u : 0u);
    const uint32_t out = ctr_load(c32 + C32_OUT);
    if (!f && ((uint64_t)out != u || u > ctl->tab_cap)) f = 8u;
    if (f) {
        ctl->fail = f;
        return;
    }
    StepLog &l = ctl->log[ctl->done % CTL_LOG];
    uint64_t lv = 0;
    for (uint32_t L = 0; L < LVL_MAX; L++) lv = max(lv, ctr_load(c64 + C64_LVL + L));
    l.n_front = n;
    l.tab_m = ctl->tab_m;
    l.children = ctr_load(c64 + C64_CHILDREN);
    l.pruned = ctr_load(c64 + C64_PRUNED);
    l.fresh = u;
    l.chw = ctr_load(c64 + C64_CHW);
    l.lv_max = lv;
    l.out = out;
    // (growth that speeds up from step to step is extrapolated once more)
    double r = n ? (double)u / (double)n : 1.0;
    if (r < 0.01) r = 0.01;
    const double trend = ctl->ratio > 0.0 && r > ctl->ratio ? r / ctl->ratio : 1.0;
    const uint64_t t = (uint64_t)(1.4 * ((double)u * r * trend * 1.1 + 4096.0));
    ctl->ratio = r;
    uint64_t m = (t + 1023) & ~1023ull;
    if (m < 1024) m = 1024;
    if (m > ctl->tab_cap) m = ctl->tab_cap;
    ctl->tab_m = m;
    ctl->n_front = u;
    ctl->row = ctl->row + 1;
    ctl->done = ctl->done + 1;
}

__global__ void k_reset_work(unsigned long long *c64) { c64[C64_WORK] = 0; }

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        p[i] = 0;
}

// Terminal states (every cell of the instance covered) add their count to
// the instance's result.
__global__ __launch_bounds__(256) void k_terminal(const uint64_t *key, const uint64_t *cnt,
                                                  uint64_t n, uint32_t row, const uint32_t *start,
                                                  const uint32_t *hgt, uint32_t inst_shift,
                                                  unsigned long long *result, const RowCtl *ctl) {
    if (ctl) {
        if (ctl->fail) return;
        n = min(n, (uint64_t)ctl->n_front);
        row = ctl->row;
    }
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint64_t k = key[i];
        const uint32_t inst = inst_shift < 64 ? (uint32_t)(k >> inst_shift) : 0u;
        if (row - start[inst] >= hgt[inst]) atomicAdd(result + inst, (unsigned long long)cnt[i]);
    }
}

// the level kernels' per-chunk counters (children written, partial rows per
// level); at a step's first chunk also the overflow flag and the chunk count
// mode bit 0: the step's first chunk (also clears the overflow flag and the
// completed-chunk count); bit 1: keep the children count, so that the next
// chunk's children follow the previous ones in the children buffer
__global__ void k_reset_levels(uint32_t *c32, unsigned long long *c64, uint32_t mode,
                               const RowCtl *ctl) {
    if (ctl && ctl->fail) return;
    const bool first = mode & 1u;
    // after an overflow the counters keep the failed chunk's sizes for the host
    if (!first && ctr_load(c32 + C32_XOVF)) return;
    if (threadIdx.x <= LVL_MAX && (threadIdx.x || !(mode & 2u))) c64[C64_CHW + threadIdx.x] = 0;
    // (children the level kernels pruned accumulate with the children kept)
    if (threadIdx.x == LVL_MAX + 1 && !(mode & 2u)) c64[C64_LPRUNED] = 0;
    if (first && threadIdx.x == 0) {
        c32[C32_XOVF] = 0;
        c32[C32_LVDONE] = 0;
    }
}

// Insert (key, cnt) into a {key, count} table of m slots, as table_insert
// does.
__device__ __forceinline__ int insert_at(unsigned long long *tab, uint64_t m, uint64_t key,
                                         uint64_t cnt) {
    uint64_t h = home_slot(key, m);
    for (uint32_t p = 0; p < 256; p++) {
        const uint64_t cur = atomicCAS(tab + 2 * h, (unsigned long long)EMPTY_KEY,
                                       (unsigned long long)key);
        if (cur == EMPTY_KEY || cur == key) {
            atomicAdd(tab + 2 * h + 1, (unsigned long long)cnt);
            return cur == EMPTY_KEY ? 1 : 0;
        }
        h = next_slot(h, m);
    }
    return -1;
}

__global__ __launch_bounds__(256) void k_rehash_table(const unsigned long long *old_tab,
                                                      uint64_t m_old, unsigned long long *new_tab,
                                                      uint64_t m_new, uint32_t *c32) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m_old;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = old_tab[2 * i];
        if (k == EMPTY_KEY) continue;
        if (insert_at(new_tab, m_new, k, old_tab[2 * i + 1]) < 0)
            atomicOr(c32 + C32_OVERFLOW, 1u);
    }
}

__global__ __launch_bounds__(256) void k_insert_spill(const uint64_t *spill, uint32_t n,
                                                      unsigned long long *tab, uint64_t m,
                                                      uint32_t *c32, unsigned long long *c64) {
    uint64_t fresh = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r = insert_at(tab, m, spill[2 * i], spill[2 * i + 1]);
        if (r < 0) atomicOr(c32 + C32_OVERFLOW, 1u);
        else fresh += (uint64_t)r;
    }
    fresh = wave_sum_u64(fresh);
    if (lane_id() == 0 && fresh) atomicAdd(c64 + C64_NEW, (unsigned long long)fresh);
}

__global__ void k_reset_donation_pair(uint32_t *c32, unsigned long long *c64, uint32_t pair) {
    c32[pair] = 0;
    c32[pair + 1] = 0;
    c64[C64_WORK] = 0;
}

__global__ void k_reset_donations(uint32_t *c32, unsigned long long *c64) {
    c32[C32_DON] = 0;
    c32[C32_DON_END] = 0;
    c64[C64_WORK] = 0;
}

// Compaction of the dedup table into the next frontier, in slot order (so the
// frontier is deterministic for a given table size): k_compact_count counts
// the occupied slots of each block's contiguous range, k_scan_blocks turns
// the counts into offsets, k_compact_write writes each range's states at its
// offset, ordered within the range by wave ballot + mbcnt and an LDS prefix
// over the block's waves.  No global atomics.
template <int SW>  // table words per slot; the key's first word leads the slot
__global__ __launch_bounds__(CMP_THREADS) void k_compact_count(
    const unsigned long long *tab_key, uint64_t m, uint64_t per_block,
    uint32_t *blk_cnt, const RowCtl *ctl) {
    __shared__ uint32_t s_w[CMP_THREADS / 64];
    if (ctl) {
        if (ctl->fail) return;
        m = min(m, (uint64_t)ctl->tab_m);
    }
    const uint64_t lo = blockIdx.x * per_block;
    const uint64_t hi = min(m, lo + per_block);
    uint32_t n = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += CMP_THREADS)
        n += tab_key[SW * i] != EMPTY_KEY;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < CMP_THREADS / 64; w++) t += s_w[w];
        blk_cnt[blockIdx.x] = t;
    }
}

// Exclusive scan of n block counts by one workgroup; total -> c32[C32_OUT].
__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t *blk_cnt, uint32_t n,
                                                      uint32_t *blk_off, uint32_t *c32,
                                                      const RowCtl *ctl) {
    __shared__ uint32_t s_part[1024];
    if (ctl && ctl->fail) return;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t lo = threadIdx.x * per, hi = min(n, lo + per);
    uint32_t t = 0;
    for (uint32_t i = lo; i < hi; i++) t += blk_cnt[i];
    s_part[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s_part[threadIdx.x] - t;  // exclusive
    for (uint32_t i = lo; i < hi; i++) {
        blk_off[i] = run;
        run += blk_cnt[i];
    }
    if (threadIdx.x == 1023) c32[C32_OUT] = s_part[1023];
}

// SW = 2: {key, count} slots (tab_cnt = tab_key + 1); SW = 4: {k0, k1,
// count, unused} slots (tab_cnt = tab_key + 2), k1 goes to out_k1.  States
// past out_cap (the host's unique count) are not written; the host compares
// c32[C32_OUT] with that count.
template <int SW>
__global__ __launch_bounds__(CMP_THREADS) void k_compact_write(
    const unsigned long long *tab_key, const unsigned long long *tab_cnt, uint64_t m,
    uint64_t per_block, const uint32_t *blk_off, uint64_t *out_key, uint64_t *out_k1,
    uint64_t *out_cnt, uint32_t *slot_uid, uint64_t out_cap, const RowCtl *ctl) {
    __shared__ uint32_t s_w[CMP_THREADS / 64];
    if (ctl) {
        if (ctl->fail) return;
        m = min(m, (uint64_t)ctl->tab_m);
    }
    const uint64_t lo = blockIdx.x * per_block;
    const uint64_t hi = min(m, lo + per_block);
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t base = blk_off[blockIdx.x];
    for (uint64_t i0 = lo; i0 < hi; i0 += CMP_THREADS) {
        const uint64_t i = i0 + threadIdx.x;
        const uint64_t k = i < hi ? tab_key[SW * i] : EMPTY_KEY;
        const bool valid = k != EMPTY_KEY;
        const uint64_t bm = __ballot(valid);
        if (lane_id() == 0) s_w[wv] = (uint32_t)__builtin_popcountll(bm);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < CMP_THREADS / 64; w++) {
            before += w < wv ? s_w[w] : 0;
            total += s_w[w];
        }
        const uint32_t o = base + before + mbcnt(bm);
        if (valid && o < out_cap) {
            out_key[o] = k;
            if (SW == 4) out_k1[o] = tab_key[SW * i + 1];
            out_cnt[o] = tab_cnt[SW * i];
            if (slot_uid) slot_uid[i] = o;
        }
        base += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_remap_edges(const uint64_t *e_slot,
                                                     const uint32_t *slot_uid,
                                                     uint32_t *e_child,
                                                     uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        e_child[i] = slot_uid[e_slot[i]];
}

__global__ __launch_bounds__(256) void k_next_size(
    const uint64_t *key, uint32_t n, uint32_t row, const uint32_t *start, const uint64_t *bw,
    uint32_t bw_stride, const uint32_t *hgt, uint32_t inst_shift,
    uint64_t occ_mask, const unsigned long long *comp, uint32_t *c32,
    unsigned long long *c64) {
    uint64_t acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        uint64_t k = key[i];
        uint32_t inst = inst_shift < 64 ? (uint32_t)(k >> inst_shift) : 0u;
        const uint32_t rr = row - start[inst];
        if (rr >= hgt[inst]) continue;
        uint64_t E = bw[(size_t)inst * bw_stride + rr] & ~(k & occ_mask);
        acc += comp[E & 0xffu];
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd(c64 + C64_NEXTSIZE, (unsigned long long)acc);
}

// Per bucket (the current row's open cells, the reference's `pos` / `ea`
// byte): records, and the row-0 completion bound of their children.
__global__ __launch_bounds__(256) void k_buckets(
    const uint64_t *key, uint64_t n, uint32_t row, const uint32_t *start, const uint64_t *bw,
    uint32_t bw_stride, const uint32_t *hgt, uint32_t inst_shift, uint64_t occ_mask,
    const unsigned long long *comp, unsigned long long *out) {
    __shared__ unsigned long long s_n[256], s_c[256];
    s_n[threadIdx.x] = 0;
    s_c[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint64_t k = key[i];
        const uint32_t inst = inst_shift < 64 ? (uint32_t)(k >> inst_shift) : 0u;
        const uint32_t rr = row - start[inst];
        if (rr >= hgt[inst]) continue;
        const uint32_t ea = (uint32_t)(bw[(size_t)inst * bw_stride + rr] & ~(k & occ_mask)) & 0xffu;
        atomicAdd(s_n + ea, 1ull);
        atomicAdd(s_c + ea, comp[ea]);
    }
    lds_barrier();
    if (s_n[threadIdx.x]) {
        atomicAdd(out + threadIdx.x, s_n[threadIdx.x]);
        atomicAdd(out + 256 + threadIdx.x, s_c[threadIdx.x]);
    }
}

// ---- first solutions on the device (enum mode) ---------------------------
// live[u] = the state's instance is in the group (terminal states).
__global__ __launch_bounds__(256) void k_fs_terminal(const uint64_t *keys, uint64_t n,
                                                     uint32_t inst_shift, const uint8_t *member,
                                                     uint8_t *live) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = inst_shift < 64 ? (uint32_t)(keys[u] >> inst_shift) : 0u;
        live[u] = member[i];
    }
}

// Backward pass over one row's edges: a parent with an edge to a live child
// is live, and best[parent] keeps the smallest descriptor of such edges in
// byte order (candidate ids in placement order).
__global__ __launch_bounds__(256) void k_fs_back(const uint32_t *par, const uint32_t *chi,
                                                 const uint64_t *desc, uint64_t n,
                                                 const uint8_t *live_child, uint8_t *live_par,
                                                 unsigned long long *best) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        if (!live_child[chi[e]]) continue;
        live_par[par[e]] = 1;
        atomicMin(best + par[e], (unsigned long long)__builtin_bswap64(desc[e]));
    }
}

// The child reached by each parent's best edge (equal descriptors from one
// parent are one completion, so the child is unique).
__global__ __launch_bounds__(256) void k_fs_pick(const uint32_t *par, const uint32_t *chi,
                                                 const uint64_t *desc, uint64_t n,
                                                 const uint8_t *live_child,
                                                 const unsigned long long *best, uint32_t *next) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x)
        if (live_child[chi[e]] && __builtin_bswap64(desc[e]) == best[par[e]])
            next[par[e]] = chi[e];
}

// root[i] = index of instance i's state in the frontier of its start step.
__global__ __launch_bounds__(256) void k_fs_roots(const uint64_t *keys, uint64_t n,
                                                  uint32_t inst_shift, const uint8_t *member,
                                                  uint32_t *root) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = inst_shift < 64 ? (uint32_t)(keys[u] >> inst_shift) : 0u;
        if (member[i]) root[i] = (uint32_t)u;
    }
}

// One thread per instance: follow the best edges from the root; out[j][r]
// is row r's descriptor, out[j][H] = 1 when the instance has a tiling.
__global__ __launch_bounds__(64) void k_fs_walk(const uint32_t *slots, uint32_t n_slots,
                                                const uint32_t *root, const uint8_t *live0,
                                                const unsigned long long *const *best,
                                                const uint32_t *const *next, uint32_t H,
                                                uint64_t *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_slots) return;
    uint64_t *o = out + (size_t)j * (H + 1);
    uint32_t u = root[slots[j]];
    if (u == 0xffffffffu || !live0[u]) {
        o[H] = 0;
        return;
    }
    for (uint32_t r = 0; r < H; r++) {
        o[r] = __builtin_bswap64(best[r][u]);
        u = next[r][u];
    }
    o[H] = 1;
}

}  // namespace

void launch_first_solutions_row(const uint32_t *par, const uint32_t *chi, const uint64_t *desc,
                                uint64_t n, const uint8_t *live_child, uint8_t *live_par,
                                unsigned long long *best, uint32_t *next, hipStream_t st) {
    const uint32_t b = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 65536);
    hipLaunchKernelGGL(k_fs_back, dim3(b), dim3(256), 0, st, par, chi, desc, n, live_child,
                       live_par, best);
    hipLaunchKernelGGL(k_fs_pick, dim3(b), dim3(256), 0, st, par, chi, desc, n, live_child,
                       (const unsigned long long *)best, next);
}

void launch_first_solutions_ends(const uint64_t *keys_term, uint64_t n_term,
                                 const uint64_t *keys_root, uint64_t n_root, uint32_t inst_shift,
                                 const uint8_t *member, uint8_t *live_term, uint32_t *root,
                                 hipStream_t st) {
    uint32_t b = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n_term + 255) / 256, 1), 65536);
    hipLaunchKernelGGL(k_fs_terminal, dim3(b), dim3(256), 0, st, keys_term, n_term, inst_shift,
                       member, live_term);
    b = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n_root + 255) / 256, 1), 65536);
    hipLaunchKernelGGL(k_fs_roots, dim3(b), dim3(256), 0, st, keys_root, n_root, inst_shift,
                       member, root);
}

void launch_first_solutions_walk(const uint32_t *slots, uint32_t n_slots, const uint32_t *root,
                                 const uint8_t *live0, const unsigned long long *const *best,
                                 const uint32_t *const *next, uint32_t H, uint64_t *out,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_fs_walk, dim3((n_slots + 63) / 64), dim3(64), 0, st, slots, n_slots,
                       root, live0, best, next, H, out);
}

size_t extend_lds_bytes(uint32_t fit_rows, uint32_t threads, bool enum_mode) {
    return sizeof(uint64_t) * (8u * fit_rows * 256u + 2u * 8u * MAX_CAND + 4u * 256u + MAX_CAND) +
           sizeof(uint64_t) * (enum_mode ? 5u : 3u) * STAGE * (threads / 64);
}

void launch_fill_table(unsigned long long *key, unsigned long long *cnt,
                       uint64_t n, hipStream_t st, const RowCtl *ctl) {
    uint64_t blocks = (n + 255) / 256;
    // (n is a capacity under a control block: fewer blocks, each striding)
    if (blocks > (ctl ? 2048 : 65536)) blocks = ctl ? 2048 : 65536;
    if (!blocks) blocks = 1;
    hipLaunchKernelGGL(k_fill_table, dim3((uint32_t)blocks), dim3(256), 0, st, key, cnt, n, ctl);
}

void launch_reset_counters(uint32_t *c32, unsigned long long *c64, hipStream_t st,
                           const RowCtl *ctl) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c32, c64, ctl);
}

void launch_copy_u64(unsigned long long *dst, const unsigned long long *src, uint64_t n,
                     hipStream_t st, const RowCtl *ctl) {
    const uint64_t b = std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 1024);
    hipLaunchKernelGGL(k_copy_u64, dim3((uint32_t)b), dim3(256), 0, st, dst, src, n, ctl);
}

void launch_row_close(RowCtl *ctl, const uint32_t *c32, const unsigned long long *c64,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_row_close, dim3(1), dim3(64), 0, st, ctl, c32, c64);
}

void launch_reset_levels(uint32_t *c32, unsigned long long *c64, bool first, hipStream_t st,
                         const RowCtl *ctl, bool keep_children) {
    hipLaunchKernelGGL(k_reset_levels, dim3(1), dim3(64), 0, st, c32, c64,
                       (uint32_t)first | (keep_children ? 2u : 0u), ctl);
}

void launch_terminal(const uint64_t *key, const uint64_t *cnt, uint64_t n, uint32_t row,
                     const uint32_t *start, const uint32_t *hgt, uint32_t inst_shift,
                     unsigned long long *result, int n_cu, hipStream_t st, const RowCtl *ctl) {
    const uint64_t b = std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), (uint64_t)n_cu * 8);
    hipLaunchKernelGGL(k_terminal, dim3((uint32_t)b), dim3(256), 0, st, key, cnt, n, row, start,
                       hgt, inst_shift, result, ctl);
}

// Random 64-bit device-scope atomic adds, iters per thread, at slots spread
// over the whole buffer by a hash of (thread, i): the device's rate for the
// access the dedup table makes per child.
__global__ __launch_bounds__(256) void k_atomic_rate(unsigned long long *tab, uint64_t m,
                                                     uint32_t iters) {
    const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
    for (uint32_t i = 0; i < iters; i++)
        atomicAdd(tab + home_slot(tid * 1000003ull + i, m), 1ull);
}

void launch_atomic_rate(unsigned long long *tab, uint64_t m, uint32_t iters, uint32_t blocks,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_atomic_rate, dim3(blocks), dim3(256), 0, st, tab, m, iters);
}

void launch_zero_u32(uint32_t *p, uint64_t n, hipStream_t st) {
    const uint64_t b = std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 1024);
    hipLaunchKernelGGL(k_zero_u32, dim3((uint32_t)b), dim3(256), 0, st, p, n);
}

void launch_reset_work(unsigned long long *c64, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, st, c64);
}

void launch_rehash(const unsigned long long *old_tab, uint64_t m_old, unsigned long long *new_tab,
                   uint64_t m_new, const uint64_t *spill, uint32_t n_spill, uint32_t *c32,
                   unsigned long long *c64, hipStream_t st) {
    launch_fill_table(new_tab, new_tab + 1, m_new, st);
    const uint64_t b1 = std::min<uint64_t>(std::max<uint64_t>((m_old + 255) / 256, 1), 65536);
    hipLaunchKernelGGL(k_rehash_table, dim3((uint32_t)b1), dim3(256), 0, st, old_tab, m_old,
                       new_tab, m_new, c32);
    if (n_spill) {
        const uint32_t b2 = (uint32_t)std::min<uint64_t>((n_spill + 255) / 256, 65536);
        hipLaunchKernelGGL(k_insert_spill, dim3(b2), dim3(256), 0, st, spill, n_spill, new_tab,
                           m_new, c32, c64);
    }
}

void launch_reset_donation_pair(uint32_t *c32, unsigned long long *c64, uint32_t pair,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_reset_donation_pair, dim3(1), dim3(1), 0, st, c32, c64, pair);
}

void launch_reset_donations(uint32_t *c32, unsigned long long *c64, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_donations, dim3(1), dim3(1), 0, st, c32, c64);
}

template <bool ENUM, bool PRUNE, bool COUNT_ONLY, int FR>
static void launch_extend_fr(const ExtendArgs &a, uint32_t threads, dim3 g, size_t lds,
                             hipStream_t st) {
    switch (threads) {
        case 256:
            hipLaunchKernelGGL((k_extend<ENUM, PRUNE, COUNT_ONLY, 256, FR>), g, dim3(256), lds, st,
                               a);
            break;
        case 512:
            hipLaunchKernelGGL((k_extend<ENUM, PRUNE, COUNT_ONLY, 512, FR>), g, dim3(512), lds, st,
                               a);
            break;
        default:
            hipLaunchKernelGGL((k_extend<ENUM, PRUNE, COUNT_ONLY, 1024, FR>), g, dim3(1024), lds,
                               st, a);
            break;
    }
}

// Table rows 1-3 (the range the tuner picks for pieces up to 5 rows tall)
// get kernels with the row count compiled in; more rows use the loop.
template <bool ENUM, bool PRUNE, bool COUNT_ONLY>
static void launch_extend_t(const ExtendArgs &a, uint32_t threads, dim3 g, size_t lds,
                            hipStream_t st) {
    switch (a.fit_rows) {
        case 1: launch_extend_fr<ENUM, PRUNE, COUNT_ONLY, 1>(a, threads, g, lds, st); break;
        case 2: launch_extend_fr<ENUM, PRUNE, COUNT_ONLY, 2>(a, threads, g, lds, st); break;
        case 3: launch_extend_fr<ENUM, PRUNE, COUNT_ONLY, 3>(a, threads, g, lds, st); break;
        default: launch_extend_fr<ENUM, PRUNE, COUNT_ONLY, 0>(a, threads, g, lds, st); break;
    }
}

void launch_extend(const ExtendArgs &a, bool enum_mode, bool prune, bool count_only,
                   uint32_t threads, int n_cu, hipStream_t st) {
    if (threads != 256 && threads != 512) threads = 1024;
    const bool en = enum_mode && !count_only;
    // enum mode stages more per child: fewer waves per workgroup if needed
    while (threads > 256 && extend_lds_bytes(a.fit_rows, threads, en) > 160u * 1024u) threads /= 2;
    const size_t lds = extend_lds_bytes(a.fit_rows, threads, en);
    // resident workgroups per CU: 160 KiB of LDS and 32 waves
    uint64_t per_cu = std::min<uint64_t>(160u * 1024u / lds, 32u / (threads / 64));
    if (!per_cu) per_cu = 1;
    // (a follow-up launch whose count is read on the device gets the full
    // resident grid; its waves leave at once when there is nothing to do)
    const uint64_t n_work = !a.don_in             ? ((uint64_t)a.n_in << (6 * a.split))
                            : a.n_don == ~0u ? ~0ull
                                             : (uint64_t)a.n_don;
    const uint64_t waves = n_work == ~0ull ? ~0ull / 2 : (n_work + 63) / 64;
    uint64_t blocks = (waves + threads / 64 - 1) / (threads / 64);
    blocks = std::min<uint64_t>(blocks, (uint64_t)n_cu * per_cu);
    if (!blocks) blocks = 1;
    const dim3 g((uint32_t)blocks);
    if (count_only) {
        if (prune) launch_extend_t<false, true, true>(a, threads, g, lds, st);
        else launch_extend_t<false, false, true>(a, threads, g, lds, st);
    } else if (enum_mode) {
        if (prune) launch_extend_t<true, true, false>(a, threads, g, lds, st);
        else launch_extend_t<true, false, false>(a, threads, g, lds, st);
    } else {
        if (prune) launch_extend_t<false, true, false>(a, threads, g, lds, st);
        else launch_extend_t<false, false, false>(a, threads, g, lds, st);
    }
}

void launch_compact(const unsigned long long *tab_key,
                    const unsigned long long *tab_cnt, uint64_t m,
                    uint64_t *out_key, uint64_t *out_cnt, uint64_t out_cap, uint32_t *c32,
                    uint32_t *blk, uint32_t *slot_uid, hipStream_t st, const RowCtl *ctl) {
    // blk holds 2 * COMPACT_MAX_BLOCKS words: counts, then offsets
    // (m is a capacity under a control block: at most 2048 blocks, so that
    // the blocks past the step's table end cost little)
    uint64_t per_block = 4096;
    while ((m + per_block - 1) / per_block > (ctl ? 2048u : COMPACT_MAX_BLOCKS)) per_block *= 2;
    const uint32_t nb = (uint32_t)((m + per_block - 1) / per_block);
    hipLaunchKernelGGL(k_compact_count<2>, dim3(nb), dim3(CMP_THREADS), 0, st, tab_key, m,
                       per_block, blk, ctl);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blk, nb,
                       blk + COMPACT_MAX_BLOCKS, c32, ctl);
    hipLaunchKernelGGL(k_compact_write<2>, dim3(nb), dim3(CMP_THREADS), 0, st, tab_key, tab_cnt,
                       m, per_block, blk + COMPACT_MAX_BLOCKS, out_key, nullptr, out_cnt,
                       slot_uid, out_cap, ctl);
}

void launch_fill_table4(unsigned long long *tab, uint64_t n, hipStream_t st) {
    uint64_t blocks = std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 65536);
    hipLaunchKernelGGL(k_fill_table4, dim3((uint32_t)blocks), dim3(256), 0, st, tab, n);
}

void launch_compact4(const unsigned long long *tab, uint64_t m, uint64_t *out_k0,
                     uint64_t *out_k1, uint64_t *out_cnt, uint64_t out_cap, uint32_t *c32,
                     uint32_t *blk, hipStream_t st) {
    uint64_t per_block = 2048;
    while ((m + per_block - 1) / per_block > COMPACT_MAX_BLOCKS) per_block *= 2;
    const uint32_t nb = (uint32_t)((m + per_block - 1) / per_block);
    hipLaunchKernelGGL(k_compact_count<4>, dim3(nb), dim3(CMP_THREADS), 0, st, tab, m,
                       per_block, blk, (const RowCtl *)nullptr);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blk, nb,
                       blk + COMPACT_MAX_BLOCKS, c32, (const RowCtl *)nullptr);
    hipLaunchKernelGGL(k_compact_write<4>, dim3(nb), dim3(CMP_THREADS), 0, st, tab, tab + 2, m,
                       per_block, blk + COMPACT_MAX_BLOCKS, out_k0, out_k1, out_cnt, nullptr,
                       out_cap, (const RowCtl *)nullptr);
}

void launch_remap_edges(const uint64_t *e_slot, const uint32_t *slot_uid,
                        uint32_t *e_child, uint32_t n, hipStream_t st) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (!blocks) blocks = 1;
    hipLaunchKernelGGL(k_remap_edges, dim3(blocks), dim3(256), 0, st, e_slot, slot_uid,
                       e_child, n);
}

void launch_buckets(const uint64_t *key, uint64_t n, uint32_t row, const uint32_t *start,
                    const uint64_t *bw, uint32_t bw_stride, const uint32_t *hgt,
                    uint32_t inst_shift, uint64_t occ_mask, const unsigned long long *comp,
                    unsigned long long *out, int n_cu, hipStream_t st) {
    launch_zero_u32(reinterpret_cast<uint32_t *>(out), 2 * 512, st);
    const uint64_t b = std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), (uint64_t)n_cu * 4);
    hipLaunchKernelGGL(k_buckets, dim3((uint32_t)b), dim3(256), 0, st, key, n, row, start, bw,
                       bw_stride, hgt, inst_shift, occ_mask, comp, out);
}

void launch_next_size(const uint64_t *key, uint32_t n, uint32_t row,
                      const uint32_t *start, const uint64_t *bw, uint32_t bw_stride,
                      const uint32_t *hgt, uint32_t inst_shift,
                      uint64_t occ_mask, const unsigned long long *comp,
                      uint32_t *c32, unsigned long long *c64, int n_cu,
                      hipStream_t st) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > (uint32_t)n_cu * 4) blocks = (uint32_t)n_cu * 4;
    if (!blocks) blocks = 1;
    hipLaunchKernelGGL(k_next_size, dim3(blocks), dim3(256), 0, st, key, n, row, start, bw,
                       bw_stride, hgt, inst_shift, occ_mask, comp, c32, c64);
}

}  // namespace d4
