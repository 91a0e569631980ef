// d4_device.h -- wavefront helpers and the connectivity test shared by the
// gfx950 kernels (64-lane waves).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "d4_internal.h"

namespace d4 {
namespace dev {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated slot allocation: every active lane calls it; lanes with
// pred get consecutive indices from *ctr with one atomic for the wave.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *ctr, bool pred) {
    uint64_t m = __ballot(pred);
    uint64_t act = __ballot(1);
    uint32_t leader = (uint32_t)__builtin_ctzll(act);
    uint32_t base = 0;
    if (lane_id() == leader && m)
        base = atomicAdd(ctr, (uint32_t)__builtin_popcountll(m));
    base = __shfl(base, (int)leader);
    return base + mbcnt(m);
}

// A device counter that earlier kernels maintain with atomics (item counts,
// flags): read with an agent-scope atomic load, which is coherent across the
// XCDs' L2 caches.  A plain load of such a counter can be served from a stale
// scalar-cache or L2 line: the partitioned dedup, whose first kernel reads
// the children count the level kernels just advanced, measured such stale
// reads (fewer children than written) until its reads were made atomic.
__device__ __forceinline__ uint64_t ctr_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ctr_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup barrier after LDS atomics: waits for this wave's LDS operations
// to complete (s_waitcnt lgkmcnt(0)) before s_barrier.  The compiler for
// gfx950 does not always emit that wait before a barrier (it was missing on
// the loop-exit path of a histogram kernel), and an LDS atomic still in
// flight when the other waves pass the barrier is then missed by their
// reads: measured as lost histogram counts, one wave's worth at a time.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt unconstrained
    __syncthreads();
}

// Loads of data a kernel reads exactly once (streaming passes): marked
// nontemporal, so they do not push out of the caches the partly written
// lines of the pass's own scattered output.
__device__ __forceinline__ uint64_t ld_once(const uint64_t *p) {
    return __builtin_nontemporal_load(p);
}
typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ ulonglong2 ld_once(const ulonglong2 *p) {
    const u64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u64x2_t *>(p));
    ulonglong2 r;
    r.x = v.x;
    r.y = v.y;
    return r;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 31;
    x *= 0x7fb5d329728ea185ull;
    x ^= x >> 27;
    x *= 0x81dadef4bc2dd44dull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
    uint32_t lo = __shfl((uint32_t)v, (int)src);
    uint32_t hi = __shfl((uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, (unsigned)o);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Dense output of one stream (wave-uniform state): lanes with `pred` get
// consecutive positions of the wave's current chunk; a chunk is reserved
// with one atomic on the stream's u64 counter.  Returns false when the
// buffer is exhausted (C32_XOVF is set; the host reruns the step larger).
struct Writer {
    uint64_t pos = 0, end = 0;
};

__device__ __forceinline__ bool reserve_chunk(unsigned long long *ctr, uint64_t chunk, uint64_t cap,
                                              uint32_t *c32, uint64_t &base) {
    uint64_t b = 0;
    if (lane_id() == 0) b = atomicAdd(ctr, (unsigned long long)chunk);
    b = shfl64(b, 0);
    if (b + chunk > cap) {
        if (lane_id() == 0) atomicOr(c32 + C32_XOVF, 1u);
        return false;
    }
    base = b;
    return true;
}

// Positions for the lanes in m: returns each lane's slot; `ok` false on
// overflow.  At most one chunk switch per call (a call places <= 64 items
// and chunks hold >= 64).
__device__ __forceinline__ uint64_t place(Writer &w, uint64_t m, unsigned long long *ctr,
                                          uint64_t chunk, uint64_t cap, uint32_t *c32, bool &ok) {
    const uint32_t n = (uint32_t)__builtin_popcountll(m);
    const uint32_t q = mbcnt(m);
    uint64_t slot = w.pos + q;
    if (w.pos + n > w.end) {
        uint64_t base = 0;
        if (!reserve_chunk(ctr, chunk, cap, c32, base)) {
            ok = false;
            return 0;
        }
        const uint64_t room = w.end - w.pos;  // slots left in the old chunk
        if (q >= room) slot = base + (q - room);
        w.pos = base + (n - room);
        w.end = base + chunk;
        return slot;
    }
    w.pos += n;
    return slot;
}


constexpr uint64_t COL0 = 0x0101010101010101ull;
constexpr uint64_t COL7 = 0x8080808080808080ull;

__device__ __forceinline__ uint64_t grow4(uint64_t x) {
    return x | (x << 8) | (x >> 8) | ((x << 1) & ~COL0) | ((x >> 1) & ~COL7);
}

// Connectivity pruning of a child's open window E (8 rows from the child's
// row).  Rows selected by `seed` hold no piece yet (pieces reach at most
// max_height-2 rows below the child's row), so every open region that meets
// them is left alone; every other 4-connected open region is closed -- it
// does not reach the window's last row -- and its area must be a multiple of
// the gcd of the piece sizes and at least the smallest piece, or no tiling
// can complete it.  (Shape::extend1 / connected / sml_island semantics,
// src/Shape.cpp:81-123, applied to the live board at run time instead of the
// reference's static table-time island test, src/frow.cpp:78-100.)
//
// Whether a closed region's area is allowed is a functor: AreaDiv tests it
// (a division), AreaMask looks it up in a mask of allowed areas made once
// (bit a: a >= min_size and a multiple of g; regions have at most 56 cells).
// The areas of the closed regions add up to the closed cells: when their
// count is not allowed, some region is not, and the flood fill of each
// region is skipped.
struct AreaDiv {
    uint32_t g, min_size;
    __device__ bool operator()(uint32_t a) const { return a >= min_size && a % g == 0; }
};
struct AreaMask {
    uint64_t m;
    __device__ bool operator()(uint32_t a) const { return (m >> a) & 1u; }
};

template <class AreaOk>
__device__ __forceinline__ bool viable(uint64_t E, uint64_t seed, AreaOk area_ok) {
    uint64_t x = E & seed;
    for (;;) {
        uint64_t y = grow4(x) & E;
        if (y == x) break;
        x = y;
    }
    uint64_t closed = E & ~x;
    if (!closed) return true;
    if (!area_ok((uint32_t)__builtin_popcountll(closed))) return false;
    while (closed) {
        uint64_t comp = closed & (~closed + 1ull);
        for (;;) {
            uint64_t y = grow4(comp) & closed;
            if (y == comp) break;
            comp = y;
        }
        if (!area_ok((uint32_t)__builtin_popcountll(comp))) return false;
        closed &= ~comp;
    }
    return true;
}
__device__ __forceinline__ bool viable(uint64_t E, uint64_t seed, uint32_t g, uint32_t min_size) {
    return viable(E, seed, AreaDiv{g, min_size});
}

// viable() for pieces at most 5 rows tall: closed regions can only lie in
// the first c = max_height - 1 <= 4 window rows (32 bits), and a cell there
// reaches the seed rows exactly when it reaches a cell of row c - 1 that is
// open above an open cell of row c (a path into the seed rows enters them
// through row c), so the flood fill runs on 32-bit words.  Same result as
// viable(E, rows >= c, ...).
__device__ __forceinline__ uint32_t grow4_32(uint32_t x) {
    return x | (x << 8) | (x >> 8) | ((x << 1) & ~0x01010101u) | ((x >> 1) & ~0x80808080u);
}
template <class AreaOk>
__device__ __forceinline__ bool viable32(uint64_t E, uint32_t c, AreaOk area_ok) {
    if (c == 0) return true;
    const uint32_t Ec = (uint32_t)E & (c >= 4 ? 0xffffffffu : (1u << (8 * c)) - 1u);
    const uint32_t below = (uint32_t)(E >> (8 * c)) & 0xffu;
    uint32_t x = Ec & (below << (8 * (c - 1)));
    for (;;) {
        const uint32_t y = grow4_32(x) & Ec;
        if (y == x) break;
        x = y;
    }
    uint32_t closed = Ec & ~x;
    if (!closed) return true;
    if (!area_ok((uint32_t)__builtin_popcount(closed))) return false;
    while (closed) {
        uint32_t comp = closed & (~closed + 1u);
        for (;;) {
            const uint32_t y = grow4_32(comp) & closed;
            if (y == comp) break;
            comp = y;
        }
        if (!area_ok((uint32_t)__builtin_popcount(comp))) return false;
        closed &= ~comp;
    }
    return true;
}
__device__ __forceinline__ bool viable32(uint64_t E, uint32_t c, uint32_t g, uint32_t min_size) {
    return viable32(E, c, AreaDiv{g, min_size});
}

// The dedup table interleaves each slot's key and count (16 B per slot:
// tab_key[2h], tab_cnt[2h] with tab_cnt = tab_key + 1), so an insert touches
// one cache line.  A table has any number m of slots (the host sizes it to
// its estimate, not to a power of two): a key's home slot is the high word
// of mix64(key) * m, and probes step linearly, wrapping at m.
__device__ __forceinline__ uint64_t home_slot(uint64_t key, uint64_t m) {
    return __umul64hi(mix64(key), m);
}
__device__ __forceinline__ uint64_t next_slot(uint64_t h, uint64_t m) {
    return h + 1 < m ? h + 1 : 0;
}
//
// Insert (key, cnt) into the table: returns 1 when the key is new, 0 when it
// merged into an existing slot, -1 when the probe limit was reached.  Keys
// only ever change EMPTY -> key, so the CAS's returned value decides: EMPTY
// means this lane claimed the slot, the own key means a merge.  One round
// trip per probe; the count update does not wait for its result.
__device__ __forceinline__ int table_insert(const ExtendArgs &a, uint64_t key,
                                            uint64_t cnt, uint64_t &slot) {
    uint64_t h = home_slot(key, a.tab_size);
    for (uint32_t p = 0; p < a.max_probe; p++) {
        uint64_t cur = atomicCAS(a.tab_key + 2 * h, (unsigned long long)EMPTY_KEY,
                                 (unsigned long long)key);
        if (cur == EMPTY_KEY || cur == key) {
            atomicAdd(a.tab_cnt + 2 * h, (unsigned long long)cnt);
            slot = h;
            return cur == EMPTY_KEY ? 1 : 0;
        }
        h = next_slot(h, a.tab_size);
    }
    return -1;
}

// FR > 0: the table has FR rows, known when compiling, so the lookups are
// independent LDS loads with constant offsets; FR == 0: `rows` at run time
// (a loop, one load after the other).
template <int FR>
__device__ __forceinline__ uint64_t fitmask(const uint64_t *s_fit,
                                            uint32_t rows, uint32_t c,
                                            uint64_t E) {
    if constexpr (FR > 0) {
        const uint64_t *f = s_fit + (size_t)c * FR * 256u;
        uint64_t m = f[E & 0xffu];
#pragma unroll
        for (int j = 1; j < FR; j++) m &= f[j * 256u + ((E >> (8 * j)) & 0xffu)];
        return m;
    } else {
        const uint64_t *f = s_fit + (size_t)c * rows * 256u;
        uint64_t m = f[E & 0xffu];
        for (uint32_t j = 1; j < rows; j++)
            m &= f[j * 256u + ((E >> (8 * j)) & 0xffu)];
        return m;
    }
}

// Candidates still available under the use counters `used`: counter fields
// never straddle a byte, so each byte indexes a table of the candidates its
// exhausted pieces rule out.
__device__ __forceinline__ uint64_t avail_of(const uint64_t *s_xa, uint32_t used) {
    return ~(s_xa[used & 0xffu] | s_xa[256 + ((used >> 8) & 0xffu)] |
             s_xa[512 + ((used >> 16) & 0xffu)] | s_xa[768 + (used >> 24)]);
}

// Packed placement entry: place[c][k] (shape) and the candidate's counter
// info (offset | field mask << 8 | count << 16, piece's candidate range
// [lo, hi) in bits 32..39 / 40..47), read with one 16-byte LDS load.
struct __attribute__((aligned(16))) PlaceInfo {
    uint64_t shape;
    uint64_t info;
};

__device__ __forceinline__ uint64_t piece_range(uint64_t info) {
    const uint32_t lo = (uint32_t)(info >> 32) & 0xffu, hi = (uint32_t)(info >> 40) & 0xffu;
    return (hi >= 64 ? ~0ull : (1ull << hi) - 1ull) & (~0ull << lo);
}

}  // namespace dev
}  // namespace d4
