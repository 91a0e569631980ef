// d4_device.h -- wavefront helpers and the connectivity test shared by the
// gfx950 kernels (64-lane waves).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

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
__device__ __forceinline__ bool viable(uint64_t E, uint64_t seed, uint32_t g,
                                       uint32_t min_size) {
    uint64_t x = E & seed;
    for (;;) {
        uint64_t y = grow4(x) & E;
        if (y == x) break;
        x = y;
    }
    uint64_t closed = E & ~x;
    while (closed) {
        uint64_t comp = closed & (~closed + 1ull);
        for (;;) {
            uint64_t y = grow4(comp) & closed;
            if (y == comp) break;
            comp = y;
        }
        uint32_t area = (uint32_t)__builtin_popcountll(comp);
        if (area < min_size || area % g) return false;
        closed &= ~comp;
    }
    return true;
}

}  // namespace dev
}  // namespace d4
