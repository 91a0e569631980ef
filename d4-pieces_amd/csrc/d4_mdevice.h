// d4_mdevice.h -- device helpers of the multi-library batch kernels
// (d4_multi.hip, d4_mlevel.hip): 128-bit candidate masks, the candidates that
// fit a window from a library's tables, and the two-word table insert.
#pragma once

#include "d4_internal.h"
#include "d4_device.h"

namespace d4 {
namespace mdev {

using namespace dev;

struct M2 {
    uint64_t lo, hi;
};

__device__ __forceinline__ M2 m_and(M2 a, M2 b) { return {a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ bool m_any(M2 a) { return (a.lo | a.hi) != 0; }
__device__ __forceinline__ uint32_t m_ctz(M2 a) {
    return a.lo ? (uint32_t)__builtin_ctzll(a.lo) : 64u + (uint32_t)__builtin_ctzll(a.hi);
}
__device__ __forceinline__ void m_pop(M2 &a) {
    if (a.lo) a.lo &= a.lo - 1;
    else a.hi &= a.hi - 1;
}
// bits [lo, hi) of a 128-bit mask
__device__ __forceinline__ uint64_t word_range(uint32_t lo, uint32_t hi, uint32_t base) {
    const int a = (int)max(lo, base) - (int)base, b = (int)min(hi, base + 64u) - (int)base;
    if (a >= b) return 0;
    const uint64_t top = b >= 64 ? ~0ull : (1ull << b) - 1ull;
    return top & (~0ull << a);
}
__device__ __forceinline__ M2 m_range(uint32_t lo, uint32_t hi) {
    return {word_range(lo, hi, 0), word_range(lo, hi, 64)};
}

__device__ __forceinline__ M2 mfit(const MLib *L, uint32_t c, uint64_t E) {
    const uint32_t i = 512u - (512u >> c) + ((uint32_t)(E & 0xffu) >> c);
    const uint32_t b1 = (uint32_t)(E >> 8) & 0xffu, b2 = (uint32_t)(E >> 16) & 0xffu;
    return {L->fit[i][0] & L->fitr[0][c][b1][0] & L->fitr[1][c][b2][0],
            L->fit[i][1] & L->fitr[0][c][b1][1] & L->fitr[1][c][b2][1]};
}

// Insert a two-word key.  Each word of a slot changes only EMPTY -> value,
// by CAS: a lane first claims or matches word 0, then claims or matches
// word 1; a slot whose word 1 holds another key is passed over.  Every
// decision is a function of the slot's final contents, so all lanes with
// the same key stop at the same slot, and no lane ever waits on another.
// Each word is read before it is CASed: a slot of another key is passed
// over with loads alone, and a key already present costs one atomic add.
__device__ __forceinline__ int table_insert2(const MultiArgs &a, uint64_t k0, uint64_t k1,
                                             uint64_t cnt) {
    uint64_t h = __umul64hi(mix64(k0 ^ mix64(k1)), a.tab_size);
    for (uint32_t p = 0; p < a.max_probe; p++) {
        unsigned long long *sl = a.tab + 4 * h;
        uint64_t c0 = __hip_atomic_load(sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c0 == EMPTY_KEY)
            c0 = atomicCAS(sl, (unsigned long long)EMPTY_KEY, (unsigned long long)k0);
        if (c0 == EMPTY_KEY || c0 == k0) {
            uint64_t c1 = __hip_atomic_load(sl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c1 == EMPTY_KEY)
                c1 = atomicCAS(sl + 1, (unsigned long long)EMPTY_KEY, (unsigned long long)k1);
            if (c1 == EMPTY_KEY || c1 == k1) {
                atomicAdd(sl + 2, (unsigned long long)cnt);
                return c1 == EMPTY_KEY ? 1 : 0;
            }
        }
        h = h + 1 < a.tab_size ? h + 1 : 0;
    }
    return -1;
}

}  // namespace mdev
}  // namespace d4
