// d4_part.hip -- dedup of a row step's children by hash buckets merged in LDS
// (gfx950).
//
// The level kernels leave every child of the row step, {key, count}, in the
// children buffer (c64[C64_CHW] entries, holes marked EMPTY_KEY).  A global
// open-addressing table merges them with about 1.6 device-scope atomics per
// child, and random 64-bit device atomics run at a fixed rate on this GPU
// whatever the table's size (about 24 G/s for a 32 KiB table as for 1 GiB;
// workgroup scope or one slice of the table per XCD change nothing), so that
// table is bound by its atomics, not by HBM.  Here the children are moved,
// in streaming passes, into buckets small enough for one workgroup's LDS,
// and merged there:
//
//   k_part_hist1    per tile (at least 64 Ki children): a survivor bit per
//                   child (one 64-bit word per wave; the connectivity prune
//                   runs here when the level kernels did not run it),
//                   survivors counted per coarse bucket (the top <= 8 bits of
//                   the key's hash) with LDS atomics
//   k_part_scan*    exclusive scan of the counts in bucket-major order: every
//                   (bucket, tile) pair gets its range of the output
//   k_part_scatter1 per tile: each survivor to its coarse bucket's range (LDS
//                   cursors); each block of 4 Ki {key, count} records is
//                   first sorted by bucket in LDS, so that a wave's 16-byte
//                   stores form runs of whole lines
//   k_part_hist2 /  the same within each coarse bucket, by the next bits of
//   k_part_scatter2 the hash, into a second buffer: a radix pass of two
//                   digits, so that a tile's survivors of one bucket land as
//                   a run of consecutive records (one digit of 14 bits would
//                   scatter them as lone 16-byte stores, which cost as much
//                   as random atomics)
//   k_part_merge    as many workgroups as fit at once, each looping over its
//                   share of the buckets: an LDS hash table of twice the
//                   pass's survivors (256 to 8 Ki slots, so clearing and
//                   scanning it costs in proportion to the bucket) per pass
//                   over a sub-range of the hash (a pass takes at most 6 Ki
//                   survivors; a bucket that has more is read once per pass,
//                   from the caches after the first), counts summed with LDS
//                   atomics, the unique states written densely to the next
//                   frontier with one device atomic per pass.
//
// There is no device atomic per child, and only surviving children (about a
// third on the 8x8 pentomino boards) are moved past the first pass.  Every size is derived on the
// device from c64[C64_CHW], so the chain runs without a host round trip and
// can be recorded in a graph.  A row with more children than the buffers were
// sized for, or a pass whose unique states exceed the LDS table (a single
// sub-range of the hash with more than 8 Ki of them), sets c32[C32_POVF]: the
// host then reruns the step with the device-memory table; a next frontier too
// small for the unique states is enlarged and the merge alone rerun.  The
// dedup checks itself: the counts it was given and the counts it emitted
// must have the same sum.  The set semantics are the
// reference's (src/sorter.cpp:184-207: equal states merge, here with their
// tiling counts summed); its radix dedup prototype (bmark/radix.cu:41-59)
// partitions by key bits in the same spirit.
#include "d4_internal.h"
#include "d4_device.h"

#include <algorithm>
#include <cstdlib>

namespace d4 {

namespace {

using namespace dev;

constexpr uint32_t PT = 1024;            // threads per workgroup (all part kernels)
constexpr uint64_t PT_TILE = 65536;      // children per level-1 tile, at least (default)
constexpr uint32_t PT_MAX_TILES = 2048;  // level-1 tiles, at most (larger rows get larger tiles)
constexpr uint64_t PT2_TILE = 65536;     // survivors per level-2 tile
constexpr uint32_t PB_MEAN = 8192;       // children (before the prune) per bucket, on average at most
constexpr uint32_t LP_MAX = 16;          // log2 of the most buckets
constexpr uint32_t LP1_MAX = 8;          // log2 of the most coarse (level-1) buckets
// LDS table slots (16 B each) of the merge: SL, a template argument (4096 or
// 8192); a pass takes at most 3/4 SL survivors (unless one sub-bin has more)
constexpr uint32_t P_SUB = 64;           // hash sub-bins a bucket's passes are cut from
constexpr uint32_t PS_BLOCK = 4096;      // scan entries per workgroup (4 per lane)
constexpr uint32_t PS_MAX_BLOCKS = 8192; // scan workgroups (8 per lane of the one top-level pass)
constexpr uint32_t LDS_INST = 1024;      // instances whose board windows go to LDS

static_assert(((uint64_t)PT_MAX_TILES << LP1_MAX) <= (uint64_t)PS_BLOCK * PS_MAX_BLOCKS,
              "level-1 histogram exceeds the scan");
static_assert(((2 * PART_MAX_CHILDREN / (PT2_TILE << LP1_MAX) + 2) << LP_MAX) <=
                  (uint64_t)PS_BLOCK * PS_MAX_BLOCKS,
              "level-2 histogram exceeds the scan");
static_assert(PART_MAX_CHILDREN <= 0xffffffffull, "bucket offsets are 32-bit");

// Buckets (log2: all, level 1, level 2), level-1 tiles and their size, and
// level-2 tiles per coarse bucket (room for twice the mean), for n children.
struct Shape {
    uint32_t lp, lp1, lp2, T1, T2;
    uint64_t tile;
};
__host__ __device__ inline Shape part_shape(uint64_t n, uint64_t tmin, uint32_t lpmax) {
    Shape s;
    const uint64_t want = (n + PB_MEAN - 1) / PB_MEAN;
    uint32_t l = 0;
    while (l < lpmax && (1ull << l) < want) l++;
    s.lp = l;
    s.lp1 = l < LP1_MAX ? l : LP1_MAX;
    s.lp2 = l - s.lp1;
    // level-1 tiles of tmin children, or n / PT_MAX_TILES rounded up to tmin
    const uint64_t per = (n + PT_MAX_TILES - 1) / PT_MAX_TILES;
    s.tile = per <= tmin ? tmin : (per + tmin - 1) / tmin * tmin;
    const uint64_t t = (n + s.tile - 1) / s.tile;
    s.T1 = (uint32_t)(t ? t : 1);
    s.T2 = (uint32_t)(2 * n / (PT2_TILE << s.lp1) + 2);
    return s;
}

__device__ __forceinline__ uint32_t bucket_of(uint64_t h, uint32_t lp) {
    return lp ? (uint32_t)(h >> (64 - lp)) : 0u;
}
__device__ __forceinline__ uint32_t sub_of(uint64_t h) { return (uint32_t)(h >> 32) & (P_SUB - 1); }

// Children of the step: none after a level overflow (the host reruns the
// step), and none when they exceed what the buffers hold (c32[C32_POVF]).
__device__ __forceinline__ uint64_t part_n(const PartArgs &a) {
    if (__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return 0;
    const uint64_t n = ctr_load(a.c64 + C64_CHW);
    const uint64_t m = n < a.ch_cap ? n : a.ch_cap;
    return m <= a.n_max ? m : 0;
}

__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, (unsigned)o);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Exclusive scan of one value per thread over the workgroup (PT threads);
// total receives the sum.  Uses s_w[PT / 64 + 1].
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t &total) {
    const uint32_t incl = wave_incl_scan32(v);
    const uint32_t wv = threadIdx.x >> 6;
    if (lane_id() == 63) s_w[wv] = incl;
    lds_barrier();
    if (threadIdx.x < 64) {
        const uint32_t x = threadIdx.x < PT / 64 ? s_w[threadIdx.x] : 0u;
        const uint32_t xi = wave_incl_scan32(x);
        if (threadIdx.x < PT / 64) s_w[threadIdx.x] = xi - x;
        if (threadIdx.x == PT / 64 - 1) s_w[PT / 64] = xi;
    }
    lds_barrier();
    total = s_w[PT / 64];
    const uint32_t r = s_w[wv] + incl - v;
    lds_barrier();  // (s_w is reused by the next call)
    return r;
}

// The children's board windows (row + 1 of each instance) into LDS.
template <bool LDSW>
__device__ __forceinline__ void stage_windows(const PartArgs &a, uint32_t row, uint64_t *s_win) {
    if (!LDSW) return;
    for (uint32_t i = threadIdx.x; i < a.n_inst; i += PT) {
        const uint32_t rr = row + 1 - a.start[i];
        s_win[i] = rr < a.bw_stride ? a.bw[(size_t)i * a.bw_stride + rr] : 0ull;
    }
}

// Connectivity prune of a child (as the table dedup's, d4_level.hip).
template <bool LDSW>
__device__ __forceinline__ bool child_viable(const PartArgs &a, uint32_t row, const uint64_t *s_win,
                                             uint64_t k) {
    const uint32_t inst = a.inst_shift < 64 ? (uint32_t)(k >> a.inst_shift) : 0u;
    uint64_t w;
    if (LDSW) {
        w = s_win[inst];
    } else {
        const uint32_t rr = row + 1 - a.start[inst];
        w = a.bw[(size_t)inst * a.bw_stride + rr];
    }
    const uint64_t E = w & ~(k & a.occ_mask);
    return a.closed_rows <= 4 ? viable32(E, a.closed_rows, AreaMask{a.area_ok})
                              : viable(E, a.seed_mask, AreaMask{a.area_ok});
}

// Level 1: prune, survivor bits, survivors per coarse bucket and tile.
template <bool PRUNE, bool LDSW>
__global__ __launch_bounds__(PT) void k_part_hist1(PartArgs a) {
    __shared__ uint64_t s_win[LDSW ? LDS_INST : 1];
    __shared__ uint32_t s_h[1u << LP1_MAX];
    if (a.ctl && a.ctl->fail) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t n0 = ctr_load(a.c64 + C64_CHW);
        if (!__hip_atomic_load(a.c32 + C32_XOVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if ((n0 < a.ch_cap ? n0 : a.ch_cap) > a.n_max) atomicOr(a.c32 + C32_POVF, 1u);
            // (children the level kernels pruned)
            const uint64_t lp = ctr_load(a.c64 + C64_LPRUNED);
            if (lp) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)lp);
        }
    }
    const uint64_t n = part_n(a);
    const Shape sh = part_shape(n, a.tile_min, a.lp_max);
    const uint32_t P1 = 1u << sh.lp1;
    const uint32_t row = a.ctl ? a.ctl->row : a.row;
    stage_windows<LDSW>(a, row, s_win);
    const uint32_t lane = lane_id();
    uint64_t kept = 0, pruned = 0;
    for (uint32_t t = blockIdx.x; t < sh.T1; t += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < P1; b += PT) s_h[b] = 0;
        lds_barrier();
        const uint64_t t0 = (uint64_t)t * sh.tile, t1 = min(t0 + sh.tile, n);
        for (uint64_t j0 = t0; j0 < t1; j0 += 4 * PT) {
            uint64_t k[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                const uint64_t i = j0 + u * PT + threadIdx.x;
                k[u] = i < t1 ? a.ch_key[i] : EMPTY_KEY;
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                bool live = k[u] != EMPTY_KEY;
                if (PRUNE && live && !child_viable<LDSW>(a, row, s_win, k[u])) {
                    live = false;
                    pruned++;
                }
                if (live) {
                    atomicAdd(&s_h[bucket_of(mix64(k[u]), sh.lp1)], 1u);
                    kept++;
                }
                // survivor bits: the wave's 64 lanes hold 64 consecutive
                // children (tiles and strides are multiples of 64)
                const uint64_t m = __ballot(live);
                const uint64_t i0 = j0 + u * PT + (threadIdx.x & ~63u);
                if (lane == 0 && i0 < t1) a.mask[i0 >> 6] = m;
            }
        }
        lds_barrier();
        for (uint32_t b = threadIdx.x; b < P1; b += PT) a.hist[(uint64_t)b * sh.T1 + t] = s_h[b];
        lds_barrier();
    }
    kept = wave_sum_u64(kept);
    pruned = wave_sum_u64(pruned);
    if (lane == 0) {
        if (kept) atomicAdd(a.c64 + C64_CHILDREN, (unsigned long long)kept);
        if (pruned) atomicAdd(a.c64 + C64_PRUNED, (unsigned long long)pruned);
    }
}

// Histogram entries of a level (bucket-major: bucket * tiles + tile).
template <int LEVEL>
__device__ __forceinline__ uint64_t hist_entries(const PartArgs &a) {
    const Shape sh = part_shape(part_n(a), a.tile_min, a.lp_max);
    return LEVEL == 1 ? ((uint64_t)1 << sh.lp1) * sh.T1 : ((uint64_t)1 << sh.lp) * sh.T2;
}
template <int LEVEL>
__device__ __forceinline__ uint32_t *hist_of(const PartArgs &a) {
    return LEVEL == 1 ? a.hist : a.hist2;
}

// Scan, level 1: the sum of each workgroup's PS_BLOCK entries.
template <int LEVEL>
__global__ __launch_bounds__(PT) void k_part_scan1(PartArgs a) {
    __shared__ uint32_t s_w[PT / 64 + 1];
    if (a.ctl && a.ctl->fail) return;
    const uint64_t m = hist_entries<LEVEL>(a);
    const uint32_t *h = hist_of<LEVEL>(a);
    const uint64_t b0 = (uint64_t)blockIdx.x * PS_BLOCK;
    if (b0 >= m) return;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        const uint64_t i = b0 + threadIdx.x * 4 + u;
        if (i < m) v += h[i];
    }
    uint32_t total;
    block_excl_scan(v, s_w, total);
    if (threadIdx.x == 0) a.bsum[blockIdx.x] = total;
}

// Scan, level 2 (one workgroup): exclusive scan of the block sums; the
// total (survivors) goes to bsum[PS_MAX_BLOCKS + LEVEL - 1].
template <int LEVEL>
__global__ __launch_bounds__(PT) void k_part_scan2(PartArgs a) {
    __shared__ uint32_t s_w[PT / 64 + 1];
    if (a.ctl && a.ctl->fail) return;
    const uint64_t m = hist_entries<LEVEL>(a);
    const uint32_t nb = (uint32_t)((m + PS_BLOCK - 1) / PS_BLOCK);
    uint32_t v[8], s = 0;
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
        const uint32_t i = threadIdx.x * 8 + u;
        v[u] = i < nb ? a.bsum[i] : 0u;
        s += v[u];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, s_w, total);
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
        const uint32_t i = threadIdx.x * 8 + u;
        if (i < nb) a.bsum[i] = run;
        run += v[u];
    }
    if (threadIdx.x == 0) a.bsum[PS_MAX_BLOCKS + LEVEL - 1] = total;
}

// Scan, level 3: each workgroup's entries become exclusive offsets, in place.
template <int LEVEL>
__global__ __launch_bounds__(PT) void k_part_scan3(PartArgs a) {
    __shared__ uint32_t s_w[PT / 64 + 1];
    if (a.ctl && a.ctl->fail) return;
    const uint64_t m = hist_entries<LEVEL>(a);
    uint32_t *h = hist_of<LEVEL>(a);
    const uint64_t b0 = (uint64_t)blockIdx.x * PS_BLOCK;
    if (b0 >= m) return;
    uint32_t v[4], s = 0;
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        const uint64_t i = b0 + threadIdx.x * 4 + u;
        v[u] = i < m ? h[i] : 0u;
        s += v[u];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, s_w, total) + a.bsum[blockIdx.x];
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        const uint64_t i = b0 + threadIdx.x * 4 + u;
        if (i < m) h[i] = run;
        run += v[u];
    }
}

// A workgroup's block of up to 4 * PT records, scattered to their buckets'
// runs (bucket b's next free position is s_cur[b]) through LDS: the block is
// first placed there in bucket order, so that consecutive lanes then store
// consecutive records of one run -- a wave's stores form a few runs of whole
// lines instead of 64 lone 16-byte stores to as many lines.  Called by every
// thread of the workgroup (it has barriers); P <= 256 buckets.
struct StageLds {
    ulonglong2 rec[4 * PT];
    uint32_t cnt[256], off[256];
    uint32_t w[PT / 64 + 1];
    uint8_t bkt[4 * PT];
};
static_assert((1u << LP1_MAX) <= 256 && (1u << (LP_MAX - LP1_MAX)) <= 256,
              "bucket ids of a staged block are bytes");

__device__ __forceinline__ void staged_scatter(StageLds &S, uint32_t *s_cur, uint32_t P,
                                               const ulonglong2 (&v)[4], const uint32_t (&b)[4],
                                               const bool (&live)[4], ulonglong2 *out) {
    for (uint32_t i = threadIdx.x; i < P; i += PT) S.cnt[i] = 0;
    lds_barrier();
    uint32_t r[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) r[u] = live[u] ? atomicAdd(&S.cnt[b[u]], 1u) : 0u;
    lds_barrier();
    uint32_t total;
    const uint32_t o = block_excl_scan(threadIdx.x < P ? S.cnt[threadIdx.x] : 0u, S.w, total);
    if (threadIdx.x < P) S.off[threadIdx.x] = o;
    lds_barrier();
#pragma unroll
    for (uint32_t u = 0; u < 4; u++)
        if (live[u]) {
            const uint32_t sl = S.off[b[u]] + r[u];
            S.rec[sl] = v[u];
            S.bkt[sl] = (uint8_t)b[u];
        }
    lds_barrier();
    for (uint32_t i = threadIdx.x; i < total; i += PT) {
        const uint32_t bb = S.bkt[i];
        out[s_cur[bb] + (i - S.off[bb])] = S.rec[i];
    }
    lds_barrier();
    if (threadIdx.x < P) s_cur[threadIdx.x] += S.cnt[threadIdx.x];
}

// Level 1: survivors to their coarse bucket's range of pbuf.
__global__ __launch_bounds__(PT) void k_part_scatter1(PartArgs a) {
    __shared__ uint32_t s_cur[1u << LP1_MAX];
    __shared__ StageLds S;
    if (a.ctl && a.ctl->fail) return;
    const uint64_t n = part_n(a);
    const Shape sh = part_shape(n, a.tile_min, a.lp_max);
    const uint32_t P1 = 1u << sh.lp1;
    const uint32_t lane = lane_id();
    ulonglong2 *pb = reinterpret_cast<ulonglong2 *>(a.pbuf);
    uint64_t csum = 0;
    for (uint32_t t = blockIdx.x; t < sh.T1; t += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < P1; b += PT) s_cur[b] = a.hist[(uint64_t)b * sh.T1 + t];
        lds_barrier();
        const uint64_t t0 = (uint64_t)t * sh.tile, t1 = min(t0 + sh.tile, n);
        for (uint64_t j0 = t0; j0 < t1; j0 += 4 * PT) {
            // (all loads of the block in flight before any dependent work)
            uint64_t m[4];
            ulonglong2 v[4];
            uint32_t bk[4];
            bool live[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                const uint64_t i0 = j0 + u * PT + (threadIdx.x & ~63u);
                m[u] = i0 < t1 ? ld_once(a.mask + (i0 >> 6)) : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                const uint64_t i = j0 + u * PT + threadIdx.x;
                live[u] = (m[u] >> lane) & 1u;
                v[u].x = live[u] ? ld_once(a.ch_key + i) : EMPTY_KEY;
                v[u].y = live[u] ? ld_once(a.ch_cnt + i) : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                bk[u] = live[u] ? bucket_of(mix64(v[u].x), sh.lp1) : 0u;
                csum += v[u].y;
            }
            staged_scatter(S, s_cur, P1, v, bk, live, pb);
        }
        lds_barrier();
    }
    csum = wave_sum_u64(csum);
    if (lane == 0 && csum) atomicAdd(a.c64 + C64_PSUM_IN, (unsigned long long)csum);
}

// A level-2 tile: part t2 of coarse bucket c, [r0, r1) of pbuf (uniform).
__device__ __forceinline__ void tile2(const PartArgs &a, const Shape &sh, uint32_t q, uint32_t &c,
                                      uint32_t &t2, uint32_t &r0, uint32_t &r1, bool &big) {
    c = q / sh.T2;
    t2 = q - c * sh.T2;
    const uint32_t P1 = 1u << sh.lp1;
    const uint32_t s = a.hist[(uint64_t)c * sh.T1];
    const uint32_t e = c + 1 < P1 ? a.hist[(uint64_t)(c + 1) * sh.T1] : a.bsum[PS_MAX_BLOCKS];
    big = (uint64_t)(e - s) > (uint64_t)sh.T2 * PT2_TILE;
    const uint64_t x0 = (uint64_t)s + (uint64_t)t2 * PT2_TILE;
    r0 = (uint32_t)min(x0, (uint64_t)e);
    r1 = (uint32_t)min(x0 + PT2_TILE, (uint64_t)e);
}

// Level 2: survivors per fine bucket within each coarse bucket's tiles.
__global__ __launch_bounds__(PT) void k_part_hist2(PartArgs a) {
    __shared__ uint32_t s_h[1u << (LP_MAX - LP1_MAX)];
    if (a.ctl && a.ctl->fail) return;
    const uint64_t n = part_n(a);
    const Shape sh = part_shape(n, a.tile_min, a.lp_max);
    const uint32_t P1 = 1u << sh.lp1, P2 = 1u << sh.lp2;
    const ulonglong2 *pb = reinterpret_cast<const ulonglong2 *>(a.pbuf);
    for (uint32_t q = blockIdx.x; q < P1 * sh.T2; q += gridDim.x) {
        uint32_t c, t2, r0, r1;
        bool big;
        tile2(a, sh, q, c, t2, r0, r1, big);
        if (big && t2 == 0 && threadIdx.x == 0) atomicOr(a.c32 + C32_POVF, 1u);
        if (threadIdx.x < P2) s_h[threadIdx.x] = 0;
        lds_barrier();
        for (uint32_t i0 = r0 + threadIdx.x; i0 < r1; i0 += 4 * PT) {
            uint64_t k[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                const uint32_t i = i0 + u * PT;
                k[u] = i < r1 ? pb[i].x : EMPTY_KEY;
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; u++)
                if (i0 + u * PT < r1) atomicAdd(&s_h[bucket_of(mix64(k[u]), sh.lp) & (P2 - 1)], 1u);
        }
        lds_barrier();
        if (threadIdx.x < P2)
            a.hist2[((uint64_t)c * P2 + threadIdx.x) * sh.T2 + t2] = s_h[threadIdx.x];
        lds_barrier();
    }
}

// Level 2: survivors from pbuf to their fine bucket's range of sbuf.
__global__ __launch_bounds__(PT) void k_part_scatter2(PartArgs a) {
    __shared__ uint32_t s_cur[1u << (LP_MAX - LP1_MAX)];
    __shared__ StageLds S;
    if (a.ctl && a.ctl->fail) return;
    const uint64_t n = part_n(a);
    const Shape sh = part_shape(n, a.tile_min, a.lp_max);
    const uint32_t P1 = 1u << sh.lp1, P2 = 1u << sh.lp2;
    const ulonglong2 *pb = reinterpret_cast<const ulonglong2 *>(a.pbuf);
    ulonglong2 *sb = reinterpret_cast<ulonglong2 *>(a.sbuf);
    for (uint32_t q = blockIdx.x; q < P1 * sh.T2; q += gridDim.x) {
        uint32_t c, t2, r0, r1;
        bool big;
        tile2(a, sh, q, c, t2, r0, r1, big);
        if (threadIdx.x < P2)
            s_cur[threadIdx.x] = a.hist2[((uint64_t)c * P2 + threadIdx.x) * sh.T2 + t2];
        lds_barrier();
        if (!big) {
            for (uint32_t j0 = r0; j0 < r1; j0 += 4 * PT) {
                ulonglong2 v[4];
                uint32_t bk[4];
                bool live[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; u++) {
                    const uint32_t i = j0 + u * PT + threadIdx.x;
                    live[u] = i < r1;
                    if (live[u]) v[u] = ld_once(pb + i);
                    else v[u].x = v[u].y = 0;
                }
#pragma unroll
                for (uint32_t u = 0; u < 4; u++)
                    bk[u] = live[u] ? bucket_of(mix64(v[u].x), sh.lp) & (P2 - 1) : 0u;
                staged_scatter(S, s_cur, P2, v, bk, live, sb);
            }
        }
        lds_barrier();
    }
}

// One workgroup per bucket: merge its survivors (children buffers, bucket
// order) in an LDS table, pass by pass over sub-ranges of the hash, and
// write each pass's unique states to the next frontier.
template <uint32_t PL_SLOTS>
__global__ __launch_bounds__(PT) void k_part_merge(PartArgs a) {
    constexpr uint32_t PL_FILL = PL_SLOTS / 4 * 3;
    __shared__ unsigned long long s_key[PL_SLOTS];
    __shared__ unsigned long long s_cnt[PL_SLOTS];
    __shared__ uint32_t s_bs[PT], s_be[PT];     // bounds of this workgroup's next buckets
    __shared__ uint32_t s_sub[P_SUB];
    __shared__ uint32_t s_pass[P_SUB + 2];      // boundaries, then the pass count
    __shared__ uint32_t s_pcnt[P_SUB];          // survivors per pass
    __shared__ uint32_t s_w[PT / 64 + 1];
    __shared__ unsigned long long s_base;
    if (a.ctl && a.ctl->fail) return;
    if (__hip_atomic_load(a.c32 + C32_POVF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u) return;
    const uint64_t n = part_n(a);
    const Shape sh = part_shape(n, a.tile_min, a.lp_max);
    const uint32_t P = 1u << sh.lp;
    if (blockIdx.x >= P) return;
    const uint32_t total = a.bsum[PS_MAX_BLOCKS + 1];
    const ulonglong2 *sb = reinterpret_cast<const ulonglong2 *>(a.sbuf);
    const uint32_t G = gridDim.x;
    const uint32_t mine = (P - blockIdx.x + G - 1) / G;  // buckets blockIdx.x + k * G
    uint32_t fail = 0;
    uint64_t csum = 0;
    for (uint32_t k0 = 0; k0 < mine; k0 += PT) {
        const uint32_t kn = min(PT, mine - k0);
        if (threadIdx.x < kn) {
            const uint32_t b = blockIdx.x + (k0 + threadIdx.x) * G;
            s_bs[threadIdx.x] = a.hist2[(uint64_t)b * sh.T2];
            s_be[threadIdx.x] = b + 1 < P ? a.hist2[(uint64_t)(b + 1) * sh.T2] : total;
        }
        lds_barrier();
        for (uint32_t k = 0; k < kn; k++) {
            const uint32_t s = s_bs[k], e = s_be[k];
            if (e <= s) continue;  // (uniform: no survivors in this bucket)
            uint32_t np = 1;
            if (e - s > PL_FILL) {
                // cut the bucket into passes of consecutive sub-bins, at most
                // PL_FILL survivors each
                if (threadIdx.x < P_SUB) s_sub[threadIdx.x] = 0;
                lds_barrier();
                for (uint32_t i = s + threadIdx.x; i < e; i += PT)
                    atomicAdd(&s_sub[sub_of(mix64(sb[i].x))], 1u);
                lds_barrier();
                if (threadIdx.x == 0) {
                    uint32_t q2 = 0, acc = 0;
                    s_pass[0] = 0;
                    for (uint32_t q = 0; q < P_SUB; q++) {
                        const uint32_t c = s_sub[q];
                        if (acc && acc + c > PL_FILL) {
                            s_pcnt[q2] = acc;
                            s_pass[++q2] = q;
                            acc = 0;
                        }
                        acc += c;
                    }
                    s_pcnt[q2] = acc;
                    s_pass[++q2] = P_SUB;
                    s_pass[P_SUB + 1] = q2;
                }
                lds_barrier();
                np = s_pass[P_SUB + 1];
            }
            for (uint32_t r = 0; r < np; r++) {
                const uint32_t lo = np == 1 ? 0u : s_pass[r], hi = np == 1 ? P_SUB : s_pass[r + 1];
                // a table of twice the pass's survivors (at least 256 slots):
                // clearing and scanning it costs in proportion to the bucket
                const uint32_t m = np == 1 ? e - s : s_pcnt[r];
                uint32_t ts = 256;
                while (ts < 2 * m && ts < PL_SLOTS) ts <<= 1;
                for (uint32_t i = threadIdx.x; i < ts; i += PT) {
                    s_key[i] = EMPTY_KEY;
                    s_cnt[i] = 0;
                }
                lds_barrier();
                for (uint32_t i0 = s + threadIdx.x; i0 < e; i0 += 4 * PT) {
                    uint64_t key[4], cnt[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4; u++) {
                        const uint32_t i = i0 + u * PT;
                        ulonglong2 v;
                        v.x = EMPTY_KEY;
                        v.y = 0;
                        if (i < e) v = sb[i];
                        key[u] = v.x;
                        cnt[u] = v.y;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; u++) {
                        if (i0 + u * PT >= e) continue;
                        const uint64_t h = mix64(key[u]);
                        const uint32_t q = sub_of(h);
                        if (q < lo || q >= hi) continue;
                        uint32_t sl = (uint32_t)h & (ts - 1), p = 0;
                        for (; p < ts; p++) {
                            const unsigned long long cur = atomicCAS(
                                &s_key[sl], (unsigned long long)EMPTY_KEY, (unsigned long long)key[u]);
                            if (cur == EMPTY_KEY || cur == key[u]) {
                                atomicAdd(&s_cnt[sl], (unsigned long long)cnt[u]);
                                break;
                            }
                            sl = (sl + 1) & (ts - 1);
                        }
                        if (p == ts) fail |= 1u;
                    }
                }
                lds_barrier();
                // this pass's unique states: slots q * PT + lane (consecutive
                // lanes read consecutive slots, free of LDS bank conflicts)
                constexpr uint32_t SPL = PL_SLOTS / PT;
                uint32_t live = 0;
#pragma unroll
                for (uint32_t q = 0; q < SPL; q++) {
                    const uint32_t sl = q * PT + threadIdx.x;
                    if (sl < ts && s_key[sl] != EMPTY_KEY) {
                        live |= 1u << q;
                        csum += s_cnt[sl];
                    }
                }
                uint32_t tot;
                const uint32_t off = block_excl_scan((uint32_t)__builtin_popcount(live), s_w, tot);
                if (threadIdx.x == 0) {
                    unsigned long long base = 0;
                    if (tot) {
                        base = atomicAdd(a.c64 + C64_NEW, (unsigned long long)tot);
                        if (base + tot <= a.out_cap) atomicAdd(a.c32 + C32_OUT, tot);
                    }
                    s_base = base;
                }
                lds_barrier();
                const unsigned long long base = s_base;
                if (base + tot <= a.out_cap) {
                    uint32_t o = off;
#pragma unroll
                    for (uint32_t q = 0; q < SPL; q++)
                        if (live >> q & 1u) {
                            const uint32_t sl = q * PT + threadIdx.x;
                            a.out_key[base + o] = s_key[sl];
                            a.out_cnt[base + o] = s_cnt[sl];
                            o++;
                        }
                } else {
                    fail |= 2u;  // (the host can rerun the merge into larger buffers)
                }
                lds_barrier();  // (the next pass clears the table)
            }
        }
        lds_barrier();  // (the next group of bounds overwrites s_bs / s_be)
    }
    if (fail) atomicOr(a.c32 + C32_POVF, fail);
    csum = wave_sum_u64(csum);
    if (lane_id() == 0 && csum) atomicAdd(a.c64 + C64_PSUM_OUT, (unsigned long long)csum);
}

// Before the merge runs again (into next-frontier buffers large enough):
// its counters back to zero.
__global__ void k_part_rearm(uint32_t *c32, unsigned long long *c64) {
    if (threadIdx.x == 0) {
        c64[C64_NEW] = 0;
        c64[C64_PSUM_OUT] = 0;
        c32[C32_OUT] = 0;
        c32[C32_POVF] = 0;
    }
}

// Buckets, level-1 tiles and level-2 tiles per coarse bucket, at most, of
// any row of n <= n_max children (the level-1 tile count is not monotone in
// n where tiles start to grow, but never exceeds PT_MAX_TILES nor
// n_max / tmin; the level-2 count grows with n).
void part_bounds(uint64_t n_max, uint64_t tmin, uint32_t lpmax, uint32_t &P, uint32_t &P1,
                 uint32_t &T1, uint32_t &T2) {
    const Shape sh = part_shape(n_max, tmin, lpmax);
    P = 1u << sh.lp;
    P1 = 1u << sh.lp1;
    T1 = (uint32_t)std::min<uint64_t>(PT_MAX_TILES, std::max<uint64_t>((n_max + tmin - 1) / tmin, 1));
    T2 = (uint32_t)(2 * n_max / (PT2_TILE << LP1_MAX) + 2);
}

// The merge runs as many workgroups as fit on the device at once (LDS-bound:
// one per CU with 8192-slot tables, two with 4096), each looping over its
// share of the buckets.
static void launch_merge(const PartArgs &a, uint32_t P, int n_cu, hipStream_t st) {
    const uint32_t per_cu = a.slots == 8192 ? 1u : 2u;
    const dim3 g(std::min<uint32_t>(P, per_cu * (uint32_t)n_cu));
    if (a.slots == 8192) hipLaunchKernelGGL(k_part_merge<8192>, g, dim3(PT), 0, st, a);
    else hipLaunchKernelGGL(k_part_merge<4096>, g, dim3(PT), 0, st, a);
}

}  // namespace

// The bucket shape, read from the environment at each call (a row step
// reads it once, in part_args): tests reach the merge's multi-pass and
// small-table paths on small searches.
uint32_t part_lp_max() {
    const char *v = getenv("D4_PART_LP");  // log2 of the most buckets
    const int x = v ? atoi(v) : 14;
    return (uint32_t)(x >= LP1_MAX && x <= (int)LP_MAX ? x : 14);
}

uint32_t part_slots() {
    const char *v = getenv("D4_PART_SLOTS");  // LDS table slots of the merge
    return (uint32_t)(v && atoi(v) == 4096 ? 4096 : 8192);
}

uint64_t part_tile_min() {
    const char *v = getenv("D4_PART_TILE");  // log2 of the smallest level-1 tile
    const int l = v ? atoi(v) : 0;
    return l >= 12 && l <= 24 ? (uint64_t)1 << l : PT_TILE;
}

void part_hist_entries(uint64_t ch_cap, uint64_t tile_min, uint32_t lp_max, uint64_t &h1,
                       uint64_t &h2) {
    uint32_t P, P1, T1, T2;
    part_bounds(std::min<uint64_t>(ch_cap, PART_MAX_CHILDREN), tile_min, lp_max, P, P1, T1, T2);
    h1 = (uint64_t)P1 * T1;
    h2 = (uint64_t)P * T2;
}

uint64_t part_bsum_entries() { return PS_MAX_BLOCKS + 2; }

uint64_t part_mask_words(uint64_t ch_cap) {
    return (std::min<uint64_t>(ch_cap, PART_MAX_CHILDREN) + 63) / 64;
}

void launch_part_dedup(const PartArgs &a, bool prune, int n_cu, hipStream_t st) {
    uint32_t P, P1, T1, T2;
    part_bounds(a.n_max, a.tile_min, a.lp_max, P, P1, T1, T2);
    const uint64_t m1 = (uint64_t)P1 * T1, m2 = (uint64_t)P * T2;
    const bool ldsw = prune && a.n_inst <= LDS_INST;
    const uint32_t g1 = std::min<uint32_t>(T1, 4u * (uint32_t)n_cu);
    const uint32_t g2 = std::min<uint32_t>(P1 * T2, 4u * (uint32_t)n_cu);
    const uint32_t s1 = (uint32_t)((m1 + PS_BLOCK - 1) / PS_BLOCK);
    const uint32_t s2 = (uint32_t)((m2 + PS_BLOCK - 1) / PS_BLOCK);
    if (!prune) hipLaunchKernelGGL((k_part_hist1<false, false>), dim3(g1), dim3(PT), 0, st, a);
    else if (ldsw) hipLaunchKernelGGL((k_part_hist1<true, true>), dim3(g1), dim3(PT), 0, st, a);
    else hipLaunchKernelGGL((k_part_hist1<true, false>), dim3(g1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan1<1>, dim3(s1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan2<1>, dim3(1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan3<1>, dim3(s1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scatter1, dim3(g1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_hist2, dim3(g2), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan1<2>, dim3(s2), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan2<2>, dim3(1), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scan3<2>, dim3(s2), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_part_scatter2, dim3(g2), dim3(PT), 0, st, a);
    launch_merge(a, P, n_cu, st);
}

void launch_part_merge_again(const PartArgs &a, int n_cu, hipStream_t st) {
    uint32_t P, P1, T1, T2;
    part_bounds(a.n_max, a.tile_min, a.lp_max, P, P1, T1, T2);
    hipLaunchKernelGGL(k_part_rearm, dim3(1), dim3(64), 0, st, a.c32, a.c64);
    launch_merge(a, P, n_cu, st);
}

}  // namespace d4
