// d4_mlevel.hip -- level-synchronous row step of the multi-library batch
// (gfx950).
//
// The single-library level kernels' structure (d4_level.hip): one launch per
// placement level over every partial row of the frontier, each wave listing
// its 64 rows' (row, candidate) pairs in LDS and processing them 64 at a
// time, one per lane, so that no lane idles on a row with few candidates.
// Here every record names its instance and, through it, its library, whose
// tables are read through the caches as in the depth-first kernel
// (d4_multi.hip): 128-bit candidate masks for the first three window rows,
// the candidates' shapes and counter fields, the candidates each counter
// byte value shows exhausted.  A row's completions are gathered per wave in
// LDS and pruned 64 at a time (the connectivity test of the child's window,
// with its library's constants); survivors are written as children, and a
// chunk's children go into the two-word table after its levels (k_mdedup),
// so a chunk whose level buffers overflow is rerun with nothing inserted
// twice.  Instances that finish at this row are counted by a pass of their
// own (k_mterminal).  The placement step is the reference searcher's
// (src/Piece.cpp:82-98): each placement covers the lowest open cell of the
// row, candidates in (piece, transform) order.
#include "d4_internal.h"
#include "d4_device.h"
#include "d4_mdevice.h"

namespace d4 {

namespace {

using namespace dev;
using namespace mdev;

constexpr int MLT = 512;          // threads per workgroup (8 waves; two workgroups per CU)
constexpr uint32_t MPCAP = 1024;  // listed (row, candidate) pairs per wave and pass

// A wave's LDS: its 64 loaded rows, the pair list, the staged children.
struct MWave {
    uint64_t E[64], BW[64], BWn[64], cnt[64], ui[64], avlo[64], avhi[64];
    uint32_t lib[64];
    uint16_t plist[MPCAP];
    uint64_t sk0[64], sk1[64], scnt[64], se[64];
};

// The candidates a library's counters leave available.
__device__ __forceinline__ M2 avail_m(const MLib *L, uint32_t used) {
    const uint64_t *x0 = L->xav[0][used & 0xffu];
    const uint64_t *x1 = L->xav[1][(used >> 8) & 0xffu];
    const uint64_t *x2 = L->xav[2][(used >> 16) & 0xffu];
    const uint64_t *x3 = L->xav[3][used >> 24];
    return {~(x0[0] | x1[0] | x2[0] | x3[0]), ~(x0[1] | x1[1] | x2[1] | x3[1])};
}

template <bool FROM_REC>
__global__ __launch_bounds__(MLT) void k_mlevel(MLevelArgs a) {
    __shared__ MWave s_w[MLT / 64];
    // the level's rows, written by the launch before (every thread reads the
    // same count): a workgroup with no batch of them leaves at once
    uint64_t n_in;
    if (FROM_REC) {
        n_in = a.n_rec;
    } else {
        n_in = ctr_load(a.c64 + C64_LVL + a.level - 1);
        if (n_in > a.in_cap) n_in = a.in_cap;
    }
    if ((uint64_t)blockIdx.x * MLT >= n_in) return;
    // a chunk before this one overflowed its level buffers (rerun from there)
    // or the table (the step is rerun): nothing this launch writes is used
    if (__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
        __hip_atomic_load(a.c32 + C32_OVERFLOW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return;

    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    MWave &W = s_w[wv];
    const uint64_t waves = (uint64_t)gridDim.x * (MLT / 64);
    const uint64_t wid = (uint64_t)blockIdx.x * (MLT / 64) + wv;
    uint64_t chunk = n_in / (waves ? waves : 1);
    chunk = chunk < 64 ? 64 : chunk > 1024 ? 1024 : chunk;
    Writer ws, wc;
    bool alive = true;
    uint64_t pruned = 0;
    unsigned long long *ctr_s = a.c64 + C64_LVL + a.level, *ctr_c = a.c64 + C64_CHW;
    uint64_t *oCnt = a.out_st, *oCw = oCnt + a.out_cap, *oUi = oCw + a.out_cap;

    // Children: gathered in the wave's staging area, tested 64 at a time
    // (k1 carries the library above bit 32 until it is written; se: the
    // child's open window, for the prune).  False when the buffer is full.
    uint32_t n_s = 0;
    auto flush = [&](uint32_t n) -> bool {
        __builtin_amdgcn_wave_barrier();
        bool keep = false;
        uint64_t k0 = 0, k1 = 0, cnt = 0;
        if (lane < n) {
            k0 = W.sk0[lane];
            k1 = W.sk1[lane];
            cnt = W.scnt[lane];
            keep = true;
            if (a.prune) {
                const MLib *Lc = a.libs + (uint32_t)(k1 >> 32);
                keep = viable(W.se[lane], Lc->seed_mask, AreaDiv{Lc->gcd, Lc->min_size});
                pruned += !keep;
            }
        }
        __builtin_amdgcn_wave_barrier();
        const uint64_t km = __ballot(keep);
        if (!km) return true;
        bool okw = true;
        const uint64_t sl = place(wc, km, ctr_c, chunk, a.ch_cap, a.c32, okw);
        if (!okw) return false;
        if (keep) {
            a.ch_k0[sl] = k0;
            a.ch_k1[sl] = k1 & 0xffffffffull;
            a.ch_cnt[sl] = cnt;
        }
        return true;
    };
    auto emit_children = [&](bool e, uint64_t k0, uint64_t k1, uint64_t cnt, uint64_t ec) -> bool {
        const uint64_t m = __ballot(e);
        if (!m) return true;
        const uint32_t nn = (uint32_t)__builtin_popcountll(m);
        const uint32_t slot = n_s + mbcnt(m);
        __builtin_amdgcn_wave_barrier();
        if (e && slot < 64) {
            W.sk0[slot] = k0;
            W.sk1[slot] = k1;
            W.scnt[slot] = cnt;
            W.se[slot] = ec;
        }
        if (n_s + nn < 64) {
            n_s += nn;
            return true;
        }
        if (!flush(64)) return false;
        if (e && slot >= 64) {
            W.sk0[slot - 64] = k0;
            W.sk1[slot - 64] = k1;
            W.scnt[slot - 64] = cnt;
            W.se[slot - 64] = ec;
        }
        n_s = n_s + nn - 64;
        return true;
    };

    const uint64_t stride = waves * 64;
    for (uint64_t b0 = wid * 64; alive && b0 < n_in; b0 += stride) {
        const uint64_t i = b0 + lane;
        const bool valid = i < n_in;
        uint64_t E = 0, BW = 0, BWn = 0, cnt = 0;
        M2 ok{0, 0}, av{0, 0};
        uint32_t used = 0, inst = 0, lib = 0;
        bool pass = false;
        if (FROM_REC) {
            if (valid) {
                const uint64_t k0 = a.in_k0[i];
                used = (uint32_t)a.in_k1[i];
                cnt = a.in_cnt[i];
                inst = (uint32_t)(k0 >> MKEY_INST_SHIFT);
                // a finished instance (counted by the terminal pass) or one
                // that left the search has no children
                const bool live = a.row < a.hgt[inst] && !(a.dropped && a.dropped[inst]);
                if (live) {
                    lib = a.inst_lib[inst];
                    const MLib *L = a.libs + lib;
                    const uint64_t *w = a.bw + (size_t)inst * a.bw_stride + a.row;
                    BW = w[0];
                    BWn = w[1];
                    E = BW & ~(k0 & L->occ_mask);
                    if (E & 0xffu) {
                        av = avail_m(L, used);
                        ok = m_and(mfit(L, (uint32_t)__builtin_ctzll(E & 0xffu), E), av);
                    } else {
                        pass = true;  // row already covered from above: its own child
                    }
                }
            }
        } else if (valid) {
            const uint64_t *s = a.in_st;
            cnt = s[i];
            const uint64_t cw = s[a.in_cap + i];
            const uint64_t ui = s[2 * a.in_cap + i];
            used = (uint32_t)ui;
            inst = (uint32_t)(ui >> 32);
            const uint32_t open = (uint32_t)(cw >> 56);
            if (open) {
                // the window and its candidates, rebuilt as the level before
                // computed them
                lib = a.inst_lib[inst];
                const MLib *L = a.libs + lib;
                const uint64_t *w = a.bw + (size_t)inst * a.bw_stride + a.row;
                BW = w[0];
                BWn = w[1];
                const uint64_t cov = cw & ((1ull << 56) - 1);
                E = (BW & ~(cov << 8) & ~0xffull) | open;
                av = avail_m(L, used);
                ok = m_and(mfit(L, (uint32_t)__builtin_ctz(open), E), av);
            }
        }
        if (FROM_REC) {
            const uint64_t k0c = ((BW & ~E) >> 8) | ((uint64_t)inst << MKEY_INST_SHIFT);
            const uint64_t ec = (E >> 8) | (BWn & 0xff00000000000000ull);
            if (!emit_children(pass, k0c, (uint64_t)used | ((uint64_t)lib << 32), cnt, ec)) {
                alive = false;
                break;
            }
        }
        __builtin_amdgcn_wave_barrier();
        W.E[lane] = E;
        W.BW[lane] = BW;
        W.BWn[lane] = BWn;
        W.cnt[lane] = cnt;
        W.ui[lane] = (uint64_t)used | ((uint64_t)inst << 32);
        W.avlo[lane] = av.lo;
        W.avhi[lane] = av.hi;
        W.lib[lane] = lib;
        const uint32_t n = (uint32_t)(__builtin_popcountll(ok.lo) + __builtin_popcountll(ok.hi));
        const uint32_t incl = wave_incl_scan(n);
        const uint32_t excl = incl - n;
        bool pending = n != 0;
        while (alive) {
            const uint64_t pendm = __ballot(pending);
            if (!pendm) break;
            const uint32_t first = (uint32_t)__builtin_ctzll(pendm);
            const uint32_t off = __shfl(excl, (int)first);
            const bool fits = pending && incl - off <= MPCAP;
            const uint64_t fm = __ballot(fits);
            const uint32_t lastf = 63u - (uint32_t)__builtin_clzll(fm);
            const uint32_t total = __shfl(incl, (int)lastf) - off;
            __builtin_amdgcn_wave_barrier();
            if (fits) {
                uint32_t t = excl - off;
                for (uint64_t m = ok.lo; m; m &= m - 1)
                    W.plist[t++] = (uint16_t)((lane << 7) | (uint32_t)__builtin_ctzll(m));
                for (uint64_t m = ok.hi; m; m &= m - 1)
                    W.plist[t++] = (uint16_t)((lane << 7) | (64u + (uint32_t)__builtin_ctzll(m)));
            }
            __builtin_amdgcn_wave_barrier();
            pending = pending && !fits;
            for (uint32_t r0 = 0; r0 < total && alive; r0 += 64) {
                const uint32_t p = r0 + lane;
                bool emitC = false, emitS = false;
                uint64_t k0c = 0, k1c = 0, ec = 0, pc = 0, cw2 = 0, ui2 = 0;
                if (p < total) {
                    const uint32_t e = W.plist[p];
                    const uint32_t o = e >> 7, k = e & 127u;
                    const uint64_t pE = W.E[o];
                    const uint32_t pl_lib = W.lib[o];
                    const MLib *L = a.libs + pl_lib;
                    const uint32_t c = (uint32_t)__builtin_ctzll(pE & 0xffu);
                    const uint64_t shape = L->cand[k][0], info = L->cand[k][1];
                    const uint64_t pl = shape << (c - ((uint32_t)(info >> 32) & 0xffu));
                    if (!(pl & ~pE)) {
                        const uint64_t pui = W.ui[o];
                        const uint64_t pBW = W.BW[o];
                        const uint64_t E2 = pE & ~pl;
                        const uint32_t used2 = (uint32_t)pui + (1u << (info & 0xffu));
                        const uint32_t pi = (uint32_t)(pui >> 32);
                        pc = W.cnt[o];
                        if (!(E2 & 0xffu)) {
                            emitC = true;
                            k0c = ((pBW & ~E2) >> 8) | ((uint64_t)pi << MKEY_INST_SHIFT);
                            k1c = (uint64_t)used2 | ((uint64_t)pl_lib << 32);
                            ec = (E2 >> 8) | (W.BWn[o] & 0xff00000000000000ull);
                        } else {
                            M2 av2{W.avlo[o], W.avhi[o]};
                            if (((used2 >> (info & 0xffu)) & ((uint32_t)(info >> 8) & 0xffu)) ==
                                ((uint32_t)(info >> 16) & 0xffffu)) {
                                const M2 rg = m_range((uint32_t)(info >> 40) & 0xffu,
                                                      (uint32_t)(info >> 48) & 0xffu);
                                av2.lo &= ~rg.lo;
                                av2.hi &= ~rg.hi;
                            }
                            const uint32_t c2 = (uint32_t)__builtin_ctzll(E2 & 0xffu);
                            emitS = m_any(m_and(mfit(L, c2, E2), av2));
                            cw2 = ((pBW & ~E2) >> 8) | ((E2 & 0xffull) << 56);
                            ui2 = (uint64_t)used2 | ((uint64_t)pi << 32);
                        }
                    }
                }
                if (!emit_children(emitC, k0c, k1c, pc, ec)) {
                    alive = false;
                    break;
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
                        oCw[sl] = cw2;
                        oUi[sl] = ui2;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // (the children still staged)
    if (alive && n_s && !flush(n_s)) alive = false;
    if (!alive) return;  // the chunk is rerun
    // unused slots of the wave's last chunks: no open cells / empty keys
    for (uint64_t j = ws.pos + lane; j < ws.end; j += 64) oCw[j] = 0;
    for (uint64_t j = wc.pos + lane; j < wc.end; j += 64) a.ch_k0[j] = EMPTY_KEY;
    pruned = wave_sum_u64(pruned);
    if (lane == 0 && pruned) atomicAdd(a.c64 + C64_LPRUNED, (unsigned long long)pruned);
}

// Terminal records: instances that finish at this row add the record's count
// to their result and, when kept, write it to the sink (as the depth-first
// kernel's absorb step does).
__global__ __launch_bounds__(256) void k_mterminal(MultiArgs a) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < a.n_in; i += (uint64_t)gridDim.x * 256) {
        const uint64_t k0 = a.in_k0[i];
        const uint32_t inst = (uint32_t)(k0 >> MKEY_INST_SHIFT);
        if (a.row < a.hgt[inst]) continue;
        const uint64_t cnt = a.in_cnt[i];
        atomicAdd(a.result + inst, (unsigned long long)cnt);
        if (a.sink) {
            // entries past the capacity are counted, not written: the host
            // grows the sink and reruns the step
            const uint32_t o = atomicAdd(a.c32 + C32_SINK, 1u);
            if (o < a.sink_cap) {
                a.sink[2 * (size_t)o] = (uint64_t)inst | ((a.in_k1[i] & 0xffffffffull) << 32);
                a.sink[2 * (size_t)o + 1] = cnt;
            }
        }
    }
}

// A completed chunk's children into the table (COUNT_ONLY: counted).
template <bool COUNT_ONLY>
__global__ __launch_bounds__(256) void k_mdedup(MultiArgs a, const uint64_t *ch_k0,
                                                const uint64_t *ch_k1, const uint64_t *ch_cnt,
                                                uint64_t ch_cap) {
    uint64_t n = ctr_load(a.c64 + C64_CHW);
    if (n > ch_cap) n = ch_cap;
    if (__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        n = 0;
    } else if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(a.c32 + C32_LVDONE, 1u);
        // the children the level kernels pruned count now that the chunk is
        // complete
        const uint64_t lp = ctr_load(a.c64 + C64_LPRUNED);
        if (lp) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)lp);
    }
    uint64_t kept = 0, fresh = 0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint64_t k0 = ch_k0[i];
        if (k0 == EMPTY_KEY) continue;
        kept++;
        if (COUNT_ONLY) continue;
        // a full table fails every later insert after max_probe probes: the
        // first failure stops every lane, and the host reruns the step larger
        if (__hip_atomic_load(a.c32 + C32_OVERFLOW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
        const int r = table_insert2(a, k0, ch_k1[i], ch_cnt[i]);
        if (r < 0) {
            atomicOr(a.c32 + C32_OVERFLOW, 1u);
            break;
        }
        fresh += r > 0;
    }
    kept = wave_sum_u64(kept);
    fresh = wave_sum_u64(fresh);
    if (lane_id() == 0) {
        if (kept) atomicAdd(a.c64 + C64_CHILDREN, (unsigned long long)kept);
        if (fresh) atomicAdd(a.c64 + C64_NEW, (unsigned long long)fresh);
    }
}

}  // namespace

void launch_mlevel(const MLevelArgs &a, int n_cu, hipStream_t st) {
    // two workgroups per CU; level 0 knows its record count
    uint64_t blocks = 2ull * (uint64_t)n_cu;
    if (a.level == 0) blocks = std::min<uint64_t>(blocks, ((uint64_t)a.n_rec + MLT - 1) / MLT);
    if (!blocks) blocks = 1;
    if (a.level == 0) hipLaunchKernelGGL(k_mlevel<true>, dim3((uint32_t)blocks), dim3(MLT), 0, st, a);
    else hipLaunchKernelGGL(k_mlevel<false>, dim3((uint32_t)blocks), dim3(MLT), 0, st, a);
}

void launch_mterminal(const MultiArgs &a, int n_cu, hipStream_t st) {
    const uint64_t b =
        std::min<uint64_t>(std::max<uint64_t>(((uint64_t)a.n_in + 255) / 256, 1), (uint64_t)n_cu * 8);
    hipLaunchKernelGGL(k_mterminal, dim3((uint32_t)b), dim3(256), 0, st, a);
}

void launch_mdedup(const MultiArgs &a, const uint64_t *ch_k0, const uint64_t *ch_k1,
                   const uint64_t *ch_cnt, uint64_t ch_cap, bool count_only, int n_cu,
                   hipStream_t st) {
    const dim3 g((uint32_t)n_cu * 8);
    if (count_only)
        hipLaunchKernelGGL(k_mdedup<true>, g, dim3(256), 0, st, a, ch_k0, ch_k1, ch_cnt, ch_cap);
    else
        hipLaunchKernelGGL(k_mdedup<false>, g, dim3(256), 0, st, a, ch_k0, ch_k1, ch_cnt, ch_cap);
}

}  // namespace d4
