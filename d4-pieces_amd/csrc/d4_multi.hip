// d4_multi.hip -- row step of a batch whose instances each have their own
// piece library and board width (thousands of uneven puzzles in one
// frontier).
//
// The single-library kernel (d4_kernels.hip) stages one library's tables in
// LDS and packs a state into one u64.  Here every record names its instance,
// the instance names its library, and the lane reads that library's tables
// from HBM through the caches (90 KiB per library): 128-bit candidate masks
// per (anchor column, row pattern) for the first three window rows -- one AND
// per row leaves the placements that fit there, as the single-library
// kernel's byte-sliced tables do (with the row-0 masks alone the search
// takes 3.6 times as long) -- the candidates' shapes and counter fields, and
// per counter byte value the candidates of the pieces it shows exhausted (a
// record's available candidates in four lookups).  States take two words, occupancy + instance
// and the use counters, so up to 14 hexominoes with multiplicity 2 on an
// 8-wide board still fit.  The search itself is the same depth-first
// enumeration of the row's completions with anchor = lowest open cell, in
// (piece, transform) order (src/Piece.cpp:82-98, src/Piece.inl:5-24).
#include "d4_internal.h"
#include "d4_device.h"
#include "d4_mdevice.h"

#include <algorithm>

namespace d4 {

namespace {

using namespace dev;
using namespace mdev;

constexpr uint32_t MSTAGE = 128;  // staged children per wave

// COUNT_ONLY runs the same search but only counts the surviving children
// (sizes the dedup table exactly while the frontier is small).
template <bool PRUNE, bool COUNT_ONLY, int THREADS>
__global__ __launch_bounds__(THREADS) void k_extend_multi(MultiArgs a) {
    __shared__ uint64_t s_stage[(THREADS / 64) * 4 * MSTAGE];
    const uint32_t lane = lane_id();
    uint64_t n_children = 0, n_new = 0, n_pruned = 0;

    // Work distribution as in the single-library kernel: a wave claims 64
    // work items with one atomic, loads the records they cover coalesced,
    // and hands them to idle lanes by shuffle.  With split levels fix = 1 / 2
    // a work item is a (record, first placement[, second placement]) triple
    // -- 128 / 16384 items per record -- so small frontiers fill the GPU.
    const bool don_mode = a.don_in != nullptr;
    const uint32_t fix = don_mode ? 0u : a.split;
    const uint32_t per_rec = 1u << (7 * fix);
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
    uint64_t bk0 = 0, bk1 = 0, bcnt = 0, bBW = 0, bBWn = 0;
    uint32_t bterm = 0, blib = 0;

    bool searching = false, pass = false;
    uint32_t inst = 0, lib = 0, used = 0, d = 0, c = 0, anc = 0, rec = 0;
    // the lane's work item owns levels >= base_d of the DFS; sib: whether
    // siblings at level base_d belong to it; steps: DFS steps on the item
    uint32_t base_d = 0, steps = 0;
    bool sib = true, no_don = false;
    uint64_t bit0 = 0, bit1 = 0;  // donated item words (don_mode)
    uint64_t cnt = 0, BW = 0, BWn = 0, E = 0, desc = 0;
    M2 avail{0, 0}, ok{0, 0};
    const MLib *L = a.libs;

    // staged children: word 0 key, word 1 counters | library << 32, word 2
    // count, word 3 open window (for pruning)
    uint64_t *st_k0 = s_stage + (threadIdx.x >> 6) * (4 * MSTAGE);
    uint64_t *st_k1 = st_k0 + MSTAGE;
    uint64_t *st_cnt = st_k1 + MSTAGE;
    uint64_t *st_E = st_cnt + MSTAGE;
    uint32_t sc = 0;
    auto flush = [&](uint32_t n) __attribute__((always_inline)) -> uint32_t {
        __builtin_amdgcn_wave_barrier();
        const bool has = lane < n;
        const uint32_t sl = sc - n + lane;
        const uint64_t k0 = has ? st_k0[sl] : 0ull;
        const uint64_t k1 = has ? st_k1[sl] : 0ull;
        const uint64_t c2 = has ? st_cnt[sl] : 0ull;
        const uint64_t e2 = has ? st_E[sl] : 0ull;
        __builtin_amdgcn_wave_barrier();
        sc -= n;
        bool keep = has;
        uint32_t t = 0;
        if (PRUNE && has) {
            const MLib *Lc = a.libs + (uint32_t)(k1 >> 32);
            keep = viable(e2, Lc->seed_mask, Lc->gcd, Lc->min_size);
            t |= keep ? 0u : 2u;
        }
        bool failed = false;
        if (keep) {
            t |= 1u;
            if (!COUNT_ONLY) {
                const int r = table_insert2(a, k0, k1 & 0xffffffffull, c2);
                if (r < 0) failed = true;
                else t |= r ? 4u : 0u;
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
    // Hand the rest of this lane's search tree to a follow-up launch: each
    // untried candidate of the current level as its own item, and at each
    // owned level below, the candidates after the one in progress.
    auto donate = [&]() __attribute__((always_inline)) -> bool {
        uint32_t n = (uint32_t)(__builtin_popcountll(ok.lo) + __builtin_popcountll(ok.hi));
        for (uint32_t l = base_d; l < d; l++) {
            const uint32_t kl = (uint32_t)((desc >> (8 * l)) & 0xffu) - 1u;
            n += (l > base_d || sib) && kl + 1 < (uint32_t)MAX_LIB_CAND;
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
        uint32_t o = b;
        auto put = [&](uint32_t lvl, uint32_t first, bool single) __attribute__((always_inline)) {
            const uint64_t pm = lvl >= 8 ? ~0ull : (1ull << (8 * lvl)) - 1ull;
            a.don_out[2 * o] = (uint64_t)rec | ((uint64_t)first << 32) | ((uint64_t)lvl << 40) |
                               ((uint64_t)single << 48);
            a.don_out[2 * o + 1] = desc & pm;
            o++;
        };
        for (M2 m = ok; m_any(m); m_pop(m)) put(d, m_ctz(m), true);
        for (uint32_t l = base_d; l < d; l++) {
            const uint32_t kl = (uint32_t)((desc >> (8 * l)) & 0xffu) - 1u;
            if ((l > base_d || sib) && kl + 1 < (uint32_t)MAX_LIB_CAND) put(l, kl + 1, false);
        }
        return true;
    };

    for (;;) {
        uint64_t im = __ballot(!(searching || pass));
        if ((uint32_t)__builtin_popcountll(im) < a.refill && ~im) im = 0;
        while (im) {
            if (buf_pos >= buf_items) {
                if (exhausted) break;
                uint64_t base = 0;
                if (lane == 0) {
                    if (__hip_atomic_load(a.c32 + C32_OVERFLOW, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT))
                        base = n_items;
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
                buf_rec0 = (uint32_t)(base >> (7 * fix));
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
                    bk0 = a.in_k0[r0];
                    bk1 = a.in_k1[r0];
                    bcnt = a.in_cnt[r0];
                    const uint32_t bi = (uint32_t)(bk0 >> MKEY_INST_SHIFT);
                    blib = a.inst_lib[bi];
                    // 1: finished (absorb), 2: dropped from the search (skip)
                    bterm = bad ? 2u
                                : a.row >= a.hgt[bi]           ? 1u
                                : (a.dropped && a.dropped[bi]) ? 2u
                                                               : 0u;
                    const uint64_t *w = a.bw + (size_t)bi * a.bw_stride + a.row;
                    bBW = bterm ? 0 : w[0];
                    bBWn = bterm ? 0 : w[1];
                }
            }
            const bool idle = (im >> lane) & 1u;
            const uint32_t q = buf_pos + mbcnt(im);
            const uint32_t take = min((uint32_t)__builtin_popcountll(im), buf_items - buf_pos);
            buf_pos += take;
            const uint32_t src = fix ? 0u : min(q, 63u);
            const uint64_t k0 = shfl64(bk0, src);
            const uint64_t k1 = shfl64(bk1, src);
            const uint64_t kc = shfl64(bcnt, src);
            const uint64_t kBW = shfl64(bBW, src);
            const uint64_t kBWn = shfl64(bBWn, src);
            const uint32_t kterm = __shfl(bterm, (int)src);
            const uint32_t klib = __shfl(blib, (int)src);
            uint64_t it0 = 0, it1 = 0;
            if (don_mode) {
                it0 = shfl64(bit0, src);
                it1 = shfl64(bit1, src);
            }
            if (idle && q < buf_items) {
                const uint32_t sub = (uint32_t)((buf_item0 + q) & (per_rec - 1));
                const bool lead_item = sub == 0 && !don_mode;
                rec = don_mode ? (uint32_t)it0 : buf_rec0 + (fix ? 0u : q);
                steps = 0;
                no_don = false;
                base_d = 0;
                sib = fix == 0;
                cnt = kc;
                inst = (uint32_t)(k0 >> MKEY_INST_SHIFT);
                if (kterm) {
                    // terminal state of a finished instance: all cells covered
                    if (kterm == 1 && lead_item && !COUNT_ONLY) {
                        atomicAdd(a.result + inst, (unsigned long long)cnt);
                        if (a.sink) {
                            // entries past the capacity are counted, not
                            // written: the host grows the sink and reruns
                            const uint32_t o = atomicAdd(a.c32 + C32_SINK, 1u);
                            if (o < a.sink_cap) {
                                a.sink[2 * (size_t)o] = (uint64_t)inst | ((k1 & 0xffffffffull) << 32);
                                a.sink[2 * (size_t)o + 1] = cnt;
                            }
                        }
                    }
                } else {
                    lib = klib;
                    L = a.libs + lib;
                    used = (uint32_t)k1;
                    BW = kBW;
                    BWn = kBWn;
                    E = BW & ~(k0 & L->occ_mask);
                    d = 0;
                    anc = 0;
                    desc = 0;
                    if (E & 0xffu) {
                        {
                            const uint64_t *x0 = L->xav[0][used & 0xffu];
                            const uint64_t *x1 = L->xav[1][(used >> 8) & 0xffu];
                            const uint64_t *x2 = L->xav[2][(used >> 16) & 0xffu];
                            const uint64_t *x3 = L->xav[3][used >> 24];
                            avail = {~(x0[0] | x1[0] | x2[0] | x3[0]), ~(x0[1] | x1[1] | x2[1] | x3[1])};
                        }
                        c = (uint32_t)__builtin_ctzll(E & 0xffu);
                        ok = m_and(mfit(L, c, E), avail);
                        if (fix == 1) {
                            ok = m_and(ok, m_range(sub, sub + 1));
                        } else if (fix == 2) {
                            // commit the first placement k0, keep only k1 next
                            const uint32_t k0i = sub >> 7, k1i = sub & 127u;
                            ok = m_and(ok, m_range(k0i, k0i + 1));
                            if (m_any(ok)) {
                                const uint64_t shape = L->cand[k0i][0], info = L->cand[k0i][1];
                                const uint64_t pl = shape << (c - ((uint32_t)(info >> 32) & 0xffu));
                                if (pl & ~E) {
                                    ok = {0, 0};
                                } else if (!((E & ~pl) & 0xffu)) {
                                    // k0 completes the row: the k1 == 0 item
                                    // emits that child on its own
                                    if (k1i) ok = {0, 0};
                                } else {
                                    E &= ~pl;
                                    used += 1u << (info & 0xffu);
                                    if (((used >> (info & 0xffu)) & ((uint32_t)(info >> 8) & 0xffu)) ==
                                        ((uint32_t)(info >> 16) & 0xffffu)) {
                                        const M2 r = m_range((uint32_t)(info >> 40) & 0xffu,
                                                             (uint32_t)(info >> 48) & 0xffu);
                                        avail.lo &= ~r.lo;
                                        avail.hi &= ~r.hi;
                                    }
                                    desc = (uint64_t)(k0i + 1);
                                    anc = c;
                                    d = 1;
                                    base_d = 1;
                                    c = (uint32_t)__builtin_ctzll(E & 0xffu);
                                    ok = m_and(m_and(mfit(L, c, E), avail), m_range(k1i, k1i + 1));
                                }
                            }
                        }
                        if (don_mode) {
                            // replay the item's placements, then take its
                            // candidates at its level
                            const uint32_t lvl = (uint32_t)(it0 >> 40) & 0xffu;
                            const uint32_t first = (uint32_t)(it0 >> 32) & 0xffu;
                            const bool single = (it0 >> 48) & 1u;
                            for (uint32_t j = 0; j < lvl; j++) {
                                const uint32_t kj = ((uint32_t)((it1 >> (8 * j)) & 0xffu) - 1u) & 127u;
                                const uint64_t shape = L->cand[kj][0], info = L->cand[kj][1];
                                E &= ~(shape << (c - ((uint32_t)(info >> 32) & 0xffu)));
                                used += 1u << (info & 0xffu);
                                if (((used >> (info & 0xffu)) & ((uint32_t)(info >> 8) & 0xffu)) ==
                                    ((uint32_t)(info >> 16) & 0xffffu)) {
                                    const M2 r = m_range((uint32_t)(info >> 40) & 0xffu,
                                                         (uint32_t)(info >> 48) & 0xffu);
                                    avail.lo &= ~r.lo;
                                    avail.hi &= ~r.hi;
                                }
                                anc |= c << (3 * j);
                                c = (uint32_t)__builtin_ctzll(E & 0xffu);
                            }
                            desc = it1;
                            d = lvl;
                            base_d = lvl;
                            sib = !single;
                            ok = m_and(m_and(mfit(L, c, E), avail),
                                       m_range(first, single ? first + 1 : 128u));
                        }
                        searching = m_any(ok);
                    } else {
                        pass = lead_item;  // row already covered from above
                    }
                }
            }
            im = __ballot(!(searching || pass));
        }
        if (!__any(searching || pass)) break;

        // one DFS step; desc holds k+1 per level (one byte each), anc the
        // anchor column (3 bits each)
        bool emit = false;
        uint64_t E2 = 0;
        uint32_t used2 = 0;
        if (pass) {
            emit = true;
            E2 = E;
            used2 = used;
            pass = false;
        } else if (searching) {
            if (a.budget && !no_don && ++steps > a.budget && donate()) {
                searching = false;
            } else if (m_any(ok)) {
                const uint32_t k = m_ctz(ok);
                m_pop(ok);
                const uint64_t shape = L->cand[k][0], info = L->cand[k][1];
                const uint64_t pl = shape << (c - ((uint32_t)(info >> 32) & 0xffu));
                const uint32_t inc = 1u << (info & 0xffu);
                E2 = E & ~pl;
                if (pl & ~E) {
                    // a lower row of the shape is blocked
                } else if (!(E2 & 0xffu)) {
                    emit = true;
                    used2 = used + inc;
                } else {
                    E = E2;
                    used += inc;
                    if (((used >> (info & 0xffu)) & ((uint32_t)(info >> 8) & 0xffu)) ==
                        ((uint32_t)(info >> 16) & 0xffffu)) {
                        const M2 r = m_range((uint32_t)(info >> 40) & 0xffu,
                                             (uint32_t)(info >> 48) & 0xffu);
                        avail.lo &= ~r.lo;
                        avail.hi &= ~r.hi;
                    }
                    desc |= (uint64_t)(k + 1) << (8 * d);
                    anc |= c << (3 * d);
                    d++;
                    c = (uint32_t)__builtin_ctzll(E & 0xffu);
                    ok = m_and(mfit(L, c, E), avail);
                }
            } else if (d <= base_d) {
                // the levels below base_d are fixed by the work item
                searching = false;
            } else {
                d--;
                const uint32_t k = (uint32_t)((desc >> (8 * d)) & 0xffu) - 1u;
                c = (anc >> (3 * d)) & 7u;
                desc &= ~(0xffull << (8 * d));
                anc &= ~(7u << (3 * d));
                const uint64_t shape = L->cand[k][0], info = L->cand[k][1];
                E |= shape << (c - ((uint32_t)(info >> 32) & 0xffu));
                used -= 1u << (info & 0xffu);
                const M2 r = m_range((uint32_t)(info >> 40) & 0xffu, (uint32_t)(info >> 48) & 0xffu);
                avail.lo |= r.lo;
                avail.hi |= r.hi;
                ok = (d > base_d || sib) ? m_and(m_and(mfit(L, c, E), avail), m_range(k + 1, 128))
                                         : M2{0, 0};
            }
        }
        const uint64_t em = __ballot(emit);
        if (em) {
            if (emit) {
                const uint32_t sl = sc + mbcnt(em);
                st_k0[sl] = ((BW & ~E2) >> 8) | ((uint64_t)inst << MKEY_INST_SHIFT);
                st_k1[sl] = (uint64_t)used2 | ((uint64_t)lib << 32);
                st_cnt[sl] = cnt;
                st_E[sl] = (E2 >> 8) | (BWn & 0xff00000000000000ull);
            }
            sc += (uint32_t)__builtin_popcountll(em);
            if (sc >= 64) tally(flush(64));
        }
    }
    while (sc) tally(flush(min(sc, 64u)));

    n_children = wave_sum_u64(n_children);
    n_new = wave_sum_u64(n_new);
    n_pruned = wave_sum_u64(n_pruned);
    if (lane == 0) {
        if (n_children) atomicAdd(a.c64 + C64_CHILDREN, (unsigned long long)n_children);
        if (n_new) atomicAdd(a.c64 + C64_NEW, (unsigned long long)n_new);
        if (n_pruned) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)n_pruned);
    }
}

// hist[i] += records of instance i in the frontier (LDS bins per
// workgroup, one global atomic per nonzero bin).
constexpr uint32_t HIST_LDS_BINS = 8192;
__global__ __launch_bounds__(256) void k_inst_hist(const uint64_t *k0, uint64_t n, uint32_t n_inst,
                                                   unsigned long long *hist) {
    __shared__ uint32_t bins[HIST_LDS_BINS];
    const bool lds = n_inst <= HIST_LDS_BINS;
    if (lds)
        for (uint32_t i = threadIdx.x; i < n_inst; i += blockDim.x) bins[i] = 0;
    __syncthreads();
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(k0[j] >> MKEY_INST_SHIFT);
        if (lds) atomicAdd(bins + i, 1u);
        else atomicAdd(hist + i, 1ull);
    }
    lds_barrier();
    if (lds)
        for (uint32_t i = threadIdx.x; i < n_inst; i += blockDim.x)
            if (bins[i]) atomicAdd(hist + i, (unsigned long long)bins[i]);
}

template <bool PRUNE, bool COUNT_ONLY>
void launch_t(const MultiArgs &a, uint32_t threads, dim3 g, hipStream_t st) {
    switch (threads) {
        case 256:
            hipLaunchKernelGGL((k_extend_multi<PRUNE, COUNT_ONLY, 256>), g, dim3(256), 0, st, a);
            break;
        case 512:
            hipLaunchKernelGGL((k_extend_multi<PRUNE, COUNT_ONLY, 512>), g, dim3(512), 0, st, a);
            break;
        default:
            hipLaunchKernelGGL((k_extend_multi<PRUNE, COUNT_ONLY, 1024>), g, dim3(1024), 0, st,
                               a);
            break;
    }
}

}  // namespace

void launch_extend_multi(const MultiArgs &a, bool prune, bool count_only, uint32_t threads,
                         int n_cu, hipStream_t st) {
    if (threads != 256 && threads != 512) threads = 1024;
    // 32 waves per CU; the stage takes 4 KiB of LDS per wave
    const uint64_t per_cu = 32u / (threads / 64);
    // (a follow-up launch whose count is read on the device gets the full
    // resident grid; its waves leave at once when there is nothing to do)
    const uint64_t n_work = !a.don_in          ? ((uint64_t)a.n_in << (7 * a.split))
                            : a.n_don == ~0u ? ~0ull / 2
                                             : (uint64_t)a.n_don;
    const uint64_t waves = (n_work + 63) / 64;
    uint64_t blocks = (waves + threads / 64 - 1) / (threads / 64);
    blocks = std::min<uint64_t>(blocks, (uint64_t)n_cu * per_cu);
    if (!blocks) blocks = 1;
    const dim3 g((uint32_t)blocks);
    if (count_only) {
        if (prune) launch_t<true, true>(a, threads, g, st);
        else launch_t<false, true>(a, threads, g, st);
    } else {
        if (prune) launch_t<true, false>(a, threads, g, st);
        else launch_t<false, false>(a, threads, g, st);
    }
}

void launch_inst_hist(const uint64_t *k0, uint64_t n, uint32_t n_inst, unsigned long long *hist,
                      int n_cu, hipStream_t st) {
    const uint64_t blocks =
        std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), (uint64_t)n_cu * 4);
    hipLaunchKernelGGL(k_inst_hist, dim3((uint32_t)blocks), dim3(256), 0, st, k0, n, n_inst, hist);
}

}  // namespace d4
