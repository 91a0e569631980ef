// d4_expand.hip -- count-mode row step over partial-row snapshots (gfx950).
//
// k_extend (d4_kernels.hip) runs one depth-first search per lane: every
// placement is a step down, every exhausted level a step back up that
// recomputes the level's candidate mask, and a partial row with no fitting
// candidate costs a step down and a step up.  k_expand enumerates the same
// row completions without backtracking:
//
//   * a partial row is a snapshot {open window E, remaining candidates ok,
//     tiling count, occupancy of the rows below, use counters, instance};
//   * every round, each lane of a wave takes the lowest remaining candidate
//     of the snapshot it holds and places it.  A placement that completes
//     the row is a child (staged, then pruned and merged into the dedup
//     table 64 at a time, as in k_extend); one that leaves the row open
//     becomes a new snapshot with its own candidate mask, or is dropped on
//     the spot when that mask is empty (a dead end costs no further round);
//   * new snapshots go on a per-wave LIFO stack in LDS (ballot + mbcnt for
//     the slots, no atomics); a lane whose snapshot has no candidate left
//     takes the newest one from the stack, and only when the stack is empty
//     does the wave draw new records.  One record's subtree thus spreads
//     over the 64 lanes of its wave instead of running on one lane.
//
// The order in which completions are found differs from the DFS, but the
// multiset of children is the same (each child is one sequence of
// placements at the lowest open cell), so the merged table is identical.
//
// Bounded memory: a snapshot that finds the stack full is written to a
// global spill buffer (EXPAND_WORDS words) and run by a follow-up launch;
// so is a wave's whole stack once the launch's record supply is exhausted
// and the wave has run `budget` more rounds (the launch's tail).  A spill
// that finds the buffer full sets C32_XOVF; the host reruns the step with a
// larger buffer.
//
// The reference's counterparts are tiled_row_search / impl<H>
// (src/kernel.cuh:79-160), whose hot loop tests every L x R half-row pair
// per record (src/kernel.cuh:131-146).
#include "d4_internal.h"
#include "d4_device.h"

#include <algorithm>

namespace d4 {

namespace {

using namespace dev;

constexpr uint32_t XSTAGE = 128;  // staged children per wave
constexpr uint32_t XCAP = 192;    // stacked snapshots per wave
constexpr uint32_t XW = EXPAND_WORDS;
static_assert(XW == 5, "spilled snapshot layout is {E, ok, cnt, cov, used | inst << 32}");

template <bool PRUNE, int THREADS, int FR>
__global__ __launch_bounds__(THREADS) void k_expand(ExtendArgs a) {
    extern __shared__ uint64_t lds[];
    const uint32_t nfit = 8u * a.fit_rows * 256u;
    uint64_t *s_fit = lds;
    const PlaceInfo *s_pc = reinterpret_cast<const PlaceInfo *>(s_fit + nfit);
    uint64_t *s_xa = s_fit + nfit + 2 * 8 * MAX_CAND;
    uint64_t *s_rng = s_xa + 4 * 256;
    for (uint32_t i = threadIdx.x; i < nfit; i += THREADS) s_fit[i] = a.fit[i];
    for (uint32_t i = threadIdx.x; i < 2 * 8 * MAX_CAND; i += THREADS)
        s_fit[nfit + i] = a.placeinfo[i];
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += THREADS) s_xa[i] = a.xavail[i];
    for (uint32_t i = threadIdx.x; i < MAX_CAND; i += THREADS)
        s_rng[i] = piece_range(a.placeinfo[2 * i + 1]);
    __syncthreads();

    // per-wave areas: the child stage, then the snapshot stack (SoA)
    uint64_t *st_key = s_rng + MAX_CAND + (threadIdx.x >> 6) * (3 * XSTAGE + XW * XCAP);
    uint64_t *st_cnt = st_key + XSTAGE;
    uint64_t *st_E = st_cnt + XSTAGE;
    uint64_t *sk_E = st_E + XSTAGE;
    uint64_t *sk_ok = sk_E + XCAP;
    uint64_t *sk_cnt = sk_ok + XCAP;
    uint64_t *sk_cov = sk_cnt + XCAP;
    uint64_t *sk_ui = sk_cov + XCAP;

    const uint32_t lane = lane_id();
    uint64_t n_children = 0, n_new = 0, n_pruned = 0;

    // Work items: records (split levels fix = 1 / 2 give 64 / 4096 items per
    // record, as in k_extend) or, in a follow-up launch, spilled snapshots.
    const bool don_mode = a.don_in != nullptr;
    const uint32_t fix = don_mode ? 0u : a.split;
    const uint32_t per_rec = 1u << (6 * fix);
    uint32_t n_don = a.n_don;
    if (don_mode && n_don == ~0u) {
        const uint32_t dn = ctr_load(a.c32 + a.don_read), de = ctr_load(a.c32 + a.don_read + 1);
        n_don = dn <= a.don_cap ? dn : (de ? de - 1 : a.don_cap);
    }
    const uint64_t n_items = don_mode ? (uint64_t)n_don : (uint64_t)a.n_in * per_rec;
    bool exhausted = false;
    uint32_t buf_items = 0, buf_pos = 0, buf_rec0 = 0;
    uint64_t buf_item0 = 0;
    // lane j of the wave holds item j of the claimed chunk: a record (key,
    // count, board window, terminal flag) or a spilled snapshot (5 words)
    uint64_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0;

    // the snapshot this lane works on
    bool has = false;
    uint64_t E = 0, ok = 0, cnt = 0, cov = 0, avail = 0;
    uint32_t used = 0, inst = 0;
    uint32_t sp = 0;    // stacked snapshots (wave-uniform)
    uint32_t sc = 0;    // staged children (wave-uniform)
    uint32_t tail = 0;  // rounds since the record supply ran out
    bool tail_off = false, gone = false;
    uint32_t rounds = 0;

    // Count mode pipelines the table insert: a flush issues the first probe
    // of each kept child and the next flush settles it.
    bool p_on = false;
    uint64_t p_key = 0, p_cnt = 0, p_h = 0, p_cur = 0;
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
            const uint32_t s = atomicAdd(a.c32 + C32_SPILL, 1u);
            if (s < a.spill_cap) {
                a.spill[2 * s] = p_key;
                a.spill[2 * s + 1] = p_cnt;
                r = 0;
            }
        }
        if (r < 0) failed = true;
        return r > 0 ? 4u : 0u;
    };
    // Stops the wave after a failure the host answers with a rerun.
    auto stop = [&]() __attribute__((always_inline)) {
        has = false;
        sp = 0;
        exhausted = true;
        buf_pos = buf_items;
    };
    // Prunes and inserts the n newest staged children, one per lane.
    // Returns this lane's tallies: bit 0 kept, bit 1 pruned, bit 2 new.
    auto flush = [&](uint32_t n) __attribute__((always_inline)) -> uint32_t {
        __builtin_amdgcn_wave_barrier();
        const bool hs = lane < n;
        const uint32_t sl = sc - n + lane;
        const uint64_t k2 = hs ? st_key[sl] : 0ull;
        const uint64_t c2 = hs ? st_cnt[sl] : 0ull;
        const uint64_t e2 = hs ? st_E[sl] : 0ull;
        __builtin_amdgcn_wave_barrier();
        sc -= n;
        bool keep = hs;
        uint32_t t = 0;
        if (PRUNE && hs) {
            keep = viable(e2, a.seed_mask, a.gcd, a.min_size);
            t |= keep ? 0u : 2u;
        }
        bool failed = false;
        t |= settle(failed);
        if (keep) {
            t |= 1u;
            p_h = home_slot(k2, a.tab_size);
            p_cur = atomicCAS(a.tab_key + 2 * p_h, (unsigned long long)EMPTY_KEY,
                              (unsigned long long)k2);
            p_key = k2;
            p_cnt = c2;
            p_on = true;
        }
        if (__any(failed)) {
            if (failed) atomicOr(a.c32 + C32_OVERFLOW, 1u);
            stop();
        }
        return t;
    };
    auto tally = [&](uint32_t t) __attribute__((always_inline)) {
        n_children += t & 1u;
        n_pruned += (t >> 1) & 1u;
        n_new += (t >> 2) & 1u;
    };
    // Reserves n spill slots for the wave (lane 0 asks, all lanes get the
    // base); false when the buffer has no room.  Once the buffer is full no
    // wave adds to the counter again, so it cannot wrap, and the one
    // reservation crossing the end records where the valid items stop.
    auto reserve = [&](uint32_t n, uint32_t &base) __attribute__((always_inline)) -> bool {
        uint32_t b = 0, okr = 0;
        if (lane == 0) {
            if (__hip_atomic_load(a.c32 + a.don_write, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) < a.don_cap) {
                b = atomicAdd(a.c32 + a.don_write, n);
                if ((uint64_t)b + n <= (uint64_t)a.don_cap) okr = 1;
                else if (b < a.don_cap) a.c32[a.don_write + 1] = b + 1;
            }
            if (!okr) atomicAdd(a.c32 + C32_DONFULL, 1u);
        }
        base = __shfl(b, 0);
        return __shfl(okr, 0) != 0;
    };
    auto put = [&](uint32_t o, uint64_t e, uint64_t k, uint64_t c, uint64_t v, uint64_t ui)
        __attribute__((always_inline)) {
        uint64_t *d = a.don_out + (size_t)XW * o;
        d[0] = e;
        d[1] = k;
        d[2] = c;
        d[3] = v;
        d[4] = ui;
    };

    for (;;) {
        rounds++;
        // 1. idle lanes take the newest stacked snapshots
        uint64_t im = __ballot(!has);
        if (im && sp) {
            const uint32_t q = mbcnt(im);
            const uint32_t npop = min((uint32_t)__builtin_popcountll(im), sp);
            __builtin_amdgcn_wave_barrier();
            if (!has && q < npop) {
                const uint32_t i = sp - 1 - q;
                E = sk_E[i];
                ok = sk_ok[i];
                cnt = sk_cnt[i];
                cov = sk_cov[i];
                const uint64_t ui = sk_ui[i];
                used = (uint32_t)ui;
                inst = (uint32_t)(ui >> 32);
                avail = avail_of(s_xa, used);
                has = true;
            }
            __builtin_amdgcn_wave_barrier();
            sp -= npop;
            im = __ballot(!has);
        }
        // 2. with the stack empty, idle lanes draw work items, once at least
        //    a.refill of them are idle (or the whole wave is)
        if ((uint32_t)__builtin_popcountll(im) < a.refill && ~im) im = 0;
        while (im) {
            if (buf_pos >= buf_items) {
                if (exhausted) break;
                uint64_t base = 0;
                if (lane == 0) {
                    if (__hip_atomic_load(a.c32 + C32_OVERFLOW, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) |
                        __hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT))
                        base = n_items;  // the step is rerun
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
                if (don_mode) {
                    if (lane < buf_items) {
                        const uint64_t *s = a.don_in + (size_t)XW * (buf_item0 + lane);
                        b0 = s[0];
                        b1 = s[1];
                        b2 = s[2];
                        b3 = s[3];
                        b4 = s[4];
                    }
                } else {
                    const uint32_t nrec = fix ? 1u : buf_items;
                    if (lane < nrec) {
                        const uint32_t r0 = buf_rec0 + lane;
                        b0 = a.in_key[r0];
                        b1 = a.in_cnt[r0];
                        const uint32_t bi =
                            a.inst_shift < 64 ? (uint32_t)(b0 >> a.inst_shift) : 0u;
                        const uint32_t rr = a.row - a.start[bi];
                        const bool term = rr >= a.hgt[bi];
                        b3 = term;
                        b2 = term ? 0 : a.bw[(size_t)bi * a.bw_stride + rr];
                    }
                }
            }
            const bool idle = (im >> lane) & 1u;
            const uint32_t q = buf_pos + mbcnt(im);
            const uint32_t take = min((uint32_t)__builtin_popcountll(im), buf_items - buf_pos);
            buf_pos += take;
            const uint32_t src = fix ? 0u : min(q, 63u);
            const uint64_t x0 = shfl64(b0, src), x1 = shfl64(b1, src), x2 = shfl64(b2, src);
            const uint64_t x3 = shfl64(b3, src);
            uint64_t x4 = 0;
            if (don_mode) x4 = shfl64(b4, src);
            bool pass = false;
            if (idle && q < buf_items) {
                if (don_mode) {
                    E = x0;
                    ok = x1;
                    cnt = x2;
                    cov = x3;
                    used = (uint32_t)x4;
                    inst = (uint32_t)(x4 >> 32);
                    avail = avail_of(s_xa, used);
                    has = ok != 0;
                } else {
                    const uint64_t item = buf_item0 + q;
                    const uint32_t sub = (uint32_t)(item & (per_rec - 1));  // k0 << 6 | k1
                    const bool lead = sub == 0;
                    const uint64_t key = x0;
                    cnt = x1;
                    inst = a.inst_shift < 64 ? (uint32_t)(key >> a.inst_shift) : 0u;
                    used = (uint32_t)(key >> a.occ_bits) & a.used_mask;
                    if (x3) {
                        // terminal state of a finished instance: all cells covered
                        if (lead) atomicAdd(a.result + inst, (unsigned long long)cnt);
                    } else {
                        const uint64_t occ = key & a.occ_mask;
                        E = x2 & ~occ;
                        cov = occ >> 8;
                        if (E & 0xffu) {
                            avail = avail_of(s_xa, used);
                            uint32_t c = (uint32_t)__builtin_ctzll(E & 0xffu);
                            ok = fitmask<FR>(s_fit, a.fit_rows, c, E) & avail;
                            if (fix == 1) {
                                ok &= 1ull << sub;
                            } else if (fix == 2) {
                                // place k0 now, keep only k1 at the next anchor
                                const uint32_t k0 = sub >> 6, k1 = sub & 63u;
                                ok &= 1ull << k0;
                                const PlaceInfo p0 = s_pc[c * MAX_CAND + k0];
                                if (ok && (p0.shape & ~E)) ok = 0;
                                if (ok) {
                                    const uint64_t E1 = E & ~p0.shape;
                                    if (!(E1 & 0xffu)) {
                                        // k0 completes the row: the k1 == 0
                                        // item emits that child on its own
                                        ok = k1 == 0 ? ok : 0;
                                    } else {
                                        const uint32_t ci = (uint32_t)p0.info;
                                        E = E1;
                                        cov |= p0.shape >> 8;
                                        used += 1u << (ci & 0xffu);
                                        if (((used >> (ci & 0xffu)) & ((ci >> 8) & 0xffu)) ==
                                            (ci >> 16))
                                            avail &= ~s_rng[k0];
                                        c = (uint32_t)__builtin_ctzll(E & 0xffu);
                                        ok = fitmask<FR>(s_fit, a.fit_rows, c, E) & avail &
                                             (1ull << k1);
                                    }
                                }
                            }
                            has = ok != 0;
                        } else {
                            pass = lead;  // row already covered from above
                        }
                    }
                }
            }
            // a record whose row is already covered is its own child
            const uint64_t pm = __ballot(pass);
            if (pm) {
                __builtin_amdgcn_wave_barrier();
                if (pass) {
                    const uint32_t sl = sc + mbcnt(pm);
                    uint64_t key2 = cov | ((uint64_t)used << a.occ_bits);
                    if (a.inst_shift < 64) key2 |= (uint64_t)inst << a.inst_shift;
                    st_key[sl] = key2;
                    st_cnt[sl] = cnt;
                    st_E[sl] = E >> 8;
                }
                __builtin_amdgcn_wave_barrier();
                sc += (uint32_t)__builtin_popcountll(pm);
                if (sc >= 64) tally(flush(64));
            }
            im = __ballot(!has);
        }
        if (!__any(has)) break;

        // 3. the launch's tail: once the supply is gone and this wave has run
        //    `budget` more rounds, its snapshots go to a follow-up launch
        //    (a wave busy with its stack learns that from the work counter)
        if (a.budget && !gone && (rounds & 31u) == 0) {
            uint32_t g = 0;
            if (lane == 0)
                g = __hip_atomic_load(a.c64 + C64_WORK, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) >= n_items;
            gone = __shfl(g, 0) != 0;
        }
        if (a.budget && (exhausted || (gone && buf_pos >= buf_items)) && !tail_off &&
            ++tail > a.budget) {
            const uint64_t hm = __ballot(has);
            const uint32_t n = sp + (uint32_t)__builtin_popcountll(hm);
            uint32_t base = 0;
            if (reserve(n, base)) {
                __builtin_amdgcn_wave_barrier();
                for (uint32_t i = lane; i < sp; i += 64)
                    put(base + i, sk_E[i], sk_ok[i], sk_cnt[i], sk_cov[i], sk_ui[i]);
                __builtin_amdgcn_wave_barrier();
                if (has)
                    put(base + sp + mbcnt(hm), E, ok, cnt, cov,
                        (uint64_t)used | ((uint64_t)inst << 32));
                if (lane == 0) atomicAdd(a.c64 + C64_DONATED, (unsigned long long)n);
                has = false;
                sp = 0;
                continue;
            }
            tail_off = true;  // no room: this wave finishes its own work
        }

        // 4. one candidate per lane
        bool emit = false, push = false;
        uint64_t E2 = 0, cov2 = 0, ok2 = 0;
        uint32_t used2 = 0;
        if (has) {
            const uint32_t c = (uint32_t)__builtin_ctzll(E & 0xffu);
            const uint32_t k = (uint32_t)__builtin_ctzll(ok);
            ok &= ok - 1;
            const PlaceInfo pk = s_pc[c * MAX_CAND + k];
            const uint64_t pl = pk.shape;
            if (!(pl & ~E)) {
                const uint32_t ci = (uint32_t)pk.info;
                E2 = E & ~pl;
                used2 = used + (1u << (ci & 0xffu));
                cov2 = cov | (pl >> 8);
                if (!(E2 & 0xffu)) {
                    emit = true;
                } else {
                    uint64_t av2 = avail;
                    if (((used2 >> (ci & 0xffu)) & ((ci >> 8) & 0xffu)) == (ci >> 16))
                        av2 &= ~s_rng[k];
                    const uint32_t c2 = (uint32_t)__builtin_ctzll(E2 & 0xffu);
                    ok2 = fitmask<FR>(s_fit, a.fit_rows, c2, E2) & av2;
                    push = ok2 != 0;
                }
            }
            has = ok != 0;
        }
        // 5. new snapshots onto the stack; beyond its capacity, to the spill
        const uint64_t pm = __ballot(push);
        if (pm) {
            const uint32_t slot = sp + mbcnt(pm);
            const bool over = push && slot >= XCAP;
            __builtin_amdgcn_wave_barrier();
            if (push && !over) {
                sk_E[slot] = E2;
                sk_ok[slot] = ok2;
                sk_cnt[slot] = cnt;
                sk_cov[slot] = cov2;
                sk_ui[slot] = (uint64_t)used2 | ((uint64_t)inst << 32);
            }
            __builtin_amdgcn_wave_barrier();
            sp = min(sp + (uint32_t)__builtin_popcountll(pm), XCAP);
            const uint64_t om = __ballot(over);
            if (om) {
                uint32_t base = 0;
                if (reserve((uint32_t)__builtin_popcountll(om), base)) {
                    if (over)
                        put(base + mbcnt(om), E2, ok2, cnt, cov2,
                            (uint64_t)used2 | ((uint64_t)inst << 32));
                    if (lane == 0)
                        atomicAdd(a.c64 + C64_DONATED,
                                  (unsigned long long)__builtin_popcountll(om));
                } else {
                    if (lane == 0) atomicOr(a.c32 + C32_XOVF, 1u);
                    stop();
                    emit = false;
                }
            }
        }
        // 6. completed rows into the stage; 64 at a time to the table
        const uint64_t em = __ballot(emit);
        if (em) {
            __builtin_amdgcn_wave_barrier();
            if (emit) {
                const uint32_t sl = sc + mbcnt(em);
                uint64_t key2 = cov2 | ((uint64_t)used2 << a.occ_bits);
                if (a.inst_shift < 64) key2 |= (uint64_t)inst << a.inst_shift;
                st_key[sl] = key2;
                st_cnt[sl] = cnt;
                st_E[sl] = E2 >> 8;
            }
            __builtin_amdgcn_wave_barrier();
            sc += (uint32_t)__builtin_popcountll(em);
            if (sc >= 64) tally(flush(64));
        }
    }
    while (sc) tally(flush(min(sc, 64u)));
    {
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

template <bool PRUNE, int FR>
static void launch_fr(const ExtendArgs &a, uint32_t threads, dim3 g, size_t lds, hipStream_t st) {
    if (threads == 512)
        hipLaunchKernelGGL((k_expand<PRUNE, 512, FR>), g, dim3(512), lds, st, a);
    else
        hipLaunchKernelGGL((k_expand<PRUNE, 256, FR>), g, dim3(256), lds, st, a);
}

template <bool PRUNE>
static void launch_p(const ExtendArgs &a, uint32_t threads, dim3 g, size_t lds, hipStream_t st) {
    switch (a.fit_rows) {
        case 1: launch_fr<PRUNE, 1>(a, threads, g, lds, st); break;
        case 2: launch_fr<PRUNE, 2>(a, threads, g, lds, st); break;
        case 3: launch_fr<PRUNE, 3>(a, threads, g, lds, st); break;
        default: launch_fr<PRUNE, 0>(a, threads, g, lds, st); break;
    }
}

}  // namespace

size_t expand_lds_bytes(uint32_t fit_rows, uint32_t threads) {
    return sizeof(uint64_t) * (8u * fit_rows * 256u + 2u * 8u * MAX_CAND + 4u * 256u + MAX_CAND) +
           sizeof(uint64_t) * (3u * XSTAGE + XW * XCAP) * (threads / 64);
}

void launch_expand(const ExtendArgs &a, bool prune, uint32_t threads, int n_cu, hipStream_t st) {
    // 512 threads when the tables and eight waves' stacks fit in 160 KiB
    threads = threads >= 512 ? 512u : 256u;
    while (threads > 256 && expand_lds_bytes(a.fit_rows, threads) > 160u * 1024u) threads /= 2;
    const size_t lds = expand_lds_bytes(a.fit_rows, threads);
    uint64_t per_cu = std::min<uint64_t>(160u * 1024u / lds, 32u / (threads / 64));
    if (!per_cu) per_cu = 1;
    const uint64_t n_work = !a.don_in             ? ((uint64_t)a.n_in << (6 * a.split))
                            : a.n_don == ~0u ? ~0ull
                                             : (uint64_t)a.n_don;
    const uint64_t waves = n_work == ~0ull ? ~0ull / 2 : (n_work + 63) / 64;
    uint64_t blocks = (waves + threads / 64 - 1) / (threads / 64);
    blocks = std::min<uint64_t>(blocks, (uint64_t)n_cu * per_cu);
    if (!blocks) blocks = 1;
    const dim3 g((uint32_t)blocks);
    if (prune) launch_p<true>(a, threads, g, lds, st);
    else launch_p<false>(a, threads, g, lds, st);
}

}  // namespace d4
