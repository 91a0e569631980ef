// d4_tables.hip -- the multi-library placement tables built on the device.
//
// The reference builds its row tables on the host, once per process
// (compute_frow_on_cpu, src/frow.cpp:39-166: every half-row empty pattern's
// completions, sorted, then copied to the device by transfer_frow_to_gpu,
// src/frow.cu:7).  A multi-library batch here holds one table set per
// distinct library (thousands in a batch of uneven instances, 92 KB each),
// so the host sends only each library's candidates and pieces (2.4 KB) and
// one workgroup per library derives the tables:
//   fit[512 - (512 >> c) + v]  candidates anchored at column c whose row-0
//                              cells lie inside row pattern v << c,
//   fitr[t][c][b]              ... whose cells in window row t + 1 lie inside b,
//   xav[byte][v]               candidates of the pieces whose use counter
//                              lies in counter byte `byte` and is exhausted
//                              when that byte reads v,
// each entry by a direct subset test over the library's placements (the host
// builder closes marked patterns over supersets; both give every placement
// whose cells in that row are a subset of the pattern).
#include <hip/hip_runtime.h>

#include "d4_internal.h"

namespace d4 {
namespace {

__device__ __forceinline__ uint32_t width_of(uint64_t s) {
    uint64_t cols = s;
    cols |= cols >> 32;
    cols |= cols >> 16;
    cols |= cols >> 8;
    const uint32_t b = (uint32_t)(cols & 0xffu);
    return b ? 32u - (uint32_t)__builtin_clz(b) : 0u;
}

__global__ __launch_bounds__(256) void k_build_mlib(const MLibSrc *src, MLib *dst) {
    __shared__ uint64_t s_pl[8][MAX_LIB_CAND];  // the placement of candidate j at column c
    __shared__ uint64_t s_pm[MAX_PIECES][2];    // the candidates of each piece
    const MLibSrc &S = src[blockIdx.x];
    MLib &M = dst[blockIdx.x];
    const uint32_t k = S.n_cand, np = S.n_pieces, t0 = threadIdx.x;
    for (uint32_t i = t0; i < 8u * MAX_LIB_CAND; i += 256) {
        const uint32_t c = i / MAX_LIB_CAND, j = i % MAX_LIB_CAND;
        uint64_t pl = 0;
        if (j < k) {
            const uint64_t n = S.cand[j][0];
            const uint32_t mc = (uint32_t)(S.cand[j][1] >> 32) & 0xffu;
            if (c >= mc && c - mc + width_of(n) <= S.W) pl = n << (c - mc);
        }
        s_pl[c][j] = pl;
    }
    if (t0 < MAX_PIECES) {
        uint64_t m0 = 0, m1 = 0;
        if (t0 < np)
            for (uint32_t j = 0; j < k; j++)
                if (S.cpiece[j] == t0) {
                    if (j < 64) m0 |= 1ull << j;
                    else m1 |= 1ull << (j - 64);
                }
        s_pm[t0][0] = m0;
        s_pm[t0][1] = m1;
    }
    __syncthreads();
    // row 0 by anchor column; entries past column 7's block stay empty
    for (uint32_t e = t0; e < 512; e += 256) {
        uint32_t c = 0;
        while (c < 8 && e >= 512u - (512u >> (c + 1))) c++;
        uint64_t m0 = 0, m1 = 0;
        if (c < 8) {
            const uint32_t b = (e - (512u - (512u >> c))) << c;
            for (uint32_t j = 0; j < k; j++) {
                const uint64_t pl = s_pl[c][j];
                if (!pl || ((uint32_t)(pl & 0xffu) & ~b)) continue;
                if (j < 64) m0 |= 1ull << j;
                else m1 |= 1ull << (j - 64);
            }
        }
        M.fit[e][0] = m0;
        M.fit[e][1] = m1;
    }
    // window rows 1 and 2
    for (uint32_t e = t0; e < 2u * 8u * 256u; e += 256) {
        const uint32_t t = e >> 11, c = (e >> 8) & 7u, b = e & 255u;
        uint64_t m0 = 0, m1 = 0;
        for (uint32_t j = 0; j < k; j++) {
            const uint64_t pl = s_pl[c][j];
            if (!pl || ((uint32_t)(pl >> (8 * (t + 1))) & 0xffu & ~b)) continue;
            if (j < 64) m0 |= 1ull << j;
            else m1 |= 1ull << (j - 64);
        }
        M.fitr[t][c][b][0] = m0;
        M.fitr[t][c][b][1] = m1;
    }
    // exhausted counters (a piece of count 0 always is)
    for (uint32_t e = t0; e < 4u * 256u; e += 256) {
        const uint32_t byte = e >> 8, v = e & 255u;
        uint64_t m0 = 0, m1 = 0;
        for (uint32_t i = 0; i < np; i++) {
            const uint64_t pc = S.piece[i];
            const uint32_t off = (uint32_t)pc & 0xffu, fmask = (uint32_t)(pc >> 8) & 0xffu;
            const uint32_t count = (uint32_t)(pc >> 16) & 0xffffu;
            if (off / 8 != byte && !(fmask == 0 && byte == 0)) continue;
            const uint32_t f = fmask ? (v >> (off % 8)) & fmask : 0u;
            if (f < count) continue;
            m0 |= s_pm[i][0];
            m1 |= s_pm[i][1];
        }
        M.xav[byte][v][0] = m0;
        M.xav[byte][v][1] = m1;
    }
    for (uint32_t j = t0; j < MAX_LIB_CAND; j += 256) {
        M.cand[j][0] = j < k ? S.cand[j][0] : 0ull;
        M.cand[j][1] = j < k ? S.cand[j][1] : 0ull;
    }
    if (t0 < MAX_PIECES) M.piece[t0] = t0 < np ? S.piece[t0] : 0ull;
    if (t0 == 0) {
        const uint32_t occ_bits = 8 * (S.maxh > 1 ? S.maxh - 1 : 0);
        M.occ_mask = occ_bits ? (1ull << occ_bits) - 1ull : 0ull;
        M.seed_mask = S.maxh >= 2 ? ~((1ull << (8 * (S.maxh - 1))) - 1ull) : ~0ull;
        M.gcd = S.gcd;
        M.min_size = S.min_size;
        M.n_pieces = np;
        M.n_cand = k;
    }
}

}  // namespace

void launch_build_mlib(const MLibSrc *src, MLib *dst, uint32_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_build_mlib, dim3(n), dim3(256), 0, st, src, dst);
}

}  // namespace d4
