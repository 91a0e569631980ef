// d4_api.cpp -- host side of the row-by-row tiling search: placement tables,
// the frontier engine that drives the gfx950 kernels one row at a time, the
// reconstruction of solutions from back-pointers, and the C ABI (d4.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "d4_internal.h"

namespace d4 {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                              \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess)                                                  \
            throw Error(e_ == hipErrorOutOfMemory ? D4_ENOMEM : D4_EDEVICE,     \
                        std::string(#x) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class F>
static int guard(F &&f) {
    try {
        f();
        return D4_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return D4_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return D4_EINVAL;
    }
}

// --------------------------------------------------------------------------
// geometry (8-wide bitboards, bit = row*8 + col)
// --------------------------------------------------------------------------

static uint64_t normalize8(uint64_t v) {
    if (!v) return 0;
    while (!(v & 0xffull)) v >>= 8;
    while (!(v & 0x0101010101010101ull)) v >>= 1;
    return v;
}

// D4 transform t in the reference's order (Shape.cpp:28-40: identity, flip
// X, flip Y, both, then the same three after a transpose).  Written as the
// image of each input cell: flips act on the input's coordinates, then the
// transpose swaps them.
static uint64_t transform8(uint64_t s, unsigned t) {
    const bool sw = t & 4, fx = t & 1, fy = t & 2;
    uint64_t out = 0;
    for (uint64_t b = s; b; b &= b - 1) {
        unsigned i = (unsigned)__builtin_ctzll(b);
        unsigned y = i / 8, x = i % 8;
        unsigned yy = fy ? 7 - y : y, xx = fx ? 7 - x : x;
        unsigned r = sw ? xx : yy, c = sw ? yy : xx;
        out |= 1ull << (r * 8 + c);
    }
    return normalize8(out);
}

static unsigned width8(uint64_t s) {
    unsigned w = 0;
    for (unsigned c = 0; c < 8; c++)
        if (s & (0x0101010101010101ull << c)) w = c + 1;
    return w;
}

static unsigned height8(uint64_t s) {
    unsigned h = 0;
    for (unsigned r = 0; r < 8; r++)
        if (s & (0xffull << (8 * r))) h = r + 1;
    return h;
}

static unsigned bitw(uint32_t v) { return v ? 32u - (unsigned)__builtin_clz(v) : 0u; }

static uint64_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Caching device allocator: freed blocks stay with the process and are
// reused by later searches of similar size (hipMalloc / hipFree of multi-GB
// tables per row would otherwise dominate short searches).  Blocks are
// rounded to 2 MiB; a request takes the smallest cached block of at least its
// size and at most twice it.  When the device runs out, the cache is
// released and the allocation retried once.
class Pool {
    std::mutex mu_;
    std::multimap<size_t, std::pair<int, void *>> free_;  // bytes -> (device, ptr)

    void trim_locked() {
        for (auto &kv : free_) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(kv.second.first);
            (void)hipFree(kv.second.second);
            (void)hipSetDevice(cur);
        }
        free_.clear();
    }

  public:
    static size_t round(size_t bytes) {
        const size_t g = 2u << 20;
        return (std::max<size_t>(bytes, 1) + g - 1) / g * g;
    }
    void *alloc(size_t bytes) {
        bytes = round(bytes);
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto it = free_.lower_bound(bytes); it != free_.end() && it->first <= 2 * bytes; ++it)
                if (it->second.first == dev) {
                    void *p = it->second.second;
                    free_.erase(it);
                    return p;
                }
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(mu_);
            trim_locked();
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess)
            throw Error(e == hipErrorOutOfMemory ? D4_ENOMEM : D4_EDEVICE,
                        std::string("hipMalloc: ") + hipGetErrorString(e));
        return p;
    }
    void release(void *p, size_t bytes) {
        if (!p) return;
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(round(bytes), std::make_pair(dev, p));
    }
};

static Pool &pool() {
    static Pool *p = new Pool();  // never destroyed: outlives every handle
    return *p;
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    uint64_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) pool().release(p, cap * sizeof(T));
        p = nullptr;
        cap = 0;
    }
    void reserve(uint64_t n) {
        if (n <= cap) return;
        release();
        uint64_t c = Pool::round(std::max<uint64_t>(n, 64) * sizeof(T)) / sizeof(T);
        p = static_cast<T *>(pool().alloc(c * sizeof(T)));
        cap = c;
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
};

}  // namespace d4

using namespace d4;

// --------------------------------------------------------------------------
// piece library and placement tables
// --------------------------------------------------------------------------

struct d4_library {
    uint32_t n = 0, sym = 0, W = 0, maxh = 0, ncand = 0, used_bits = 0, fit_rows = 1;
    uint32_t counts[MAX_PIECES] = {}, sizes[MAX_PIECES] = {};
    uint32_t off[MAX_PIECES] = {}, fbits[MAX_PIECES] = {};
    uint8_t cpiece[MAX_CAND] = {}, ctrs[MAX_CAND] = {}, cminc[MAX_CAND] = {};
    uint64_t cnormal[MAX_CAND] = {};
    uint64_t place[8][MAX_CAND] = {};
    uint64_t placeinfo[8 * MAX_CAND * 2] = {};
    uint32_t ext_threads = 1024;  // extend workgroup size (autotuned)
    uint64_t range[MAX_CAND] = {};
    uint32_t cinfo[MAX_CAND] = {};
    uint64_t prange[MAX_PIECES] = {};
    uint64_t xavail[4 * 256] = {};
    std::vector<uint64_t> fit;  // [8][maxh][256]
    unsigned long long comp[256] = {};
    uint32_t gcd = 1, min_size = 1;
    uint32_t lds_limit = 65536;
    int device = 0;
    int n_cu = 1;
    uint64_t *d_fit = nullptr, *d_placeinfo = nullptr, *d_xavail = nullptr;
    unsigned long long *d_comp = nullptr;

    ~d4_library() {
        for (void *p : {(void *)d_fit, (void *)d_placeinfo, (void *)d_xavail, (void *)d_comp})
            if (p) (void)hipFree(p);
    }
};

// fit[c][r][b]: candidates anchored at column c whose row-r cells lie in b.
static void build_fit(d4_library &L, uint32_t fit_rows) {
    L.fit_rows = fit_rows;
    L.fit.assign((size_t)8 * fit_rows * 256, 0);
    for (uint32_t c = 0; c < 8; c++)
        for (uint32_t r = 0; r < fit_rows; r++)
            for (uint32_t b = 0; b < 256; b++) {
                uint64_t m = 0;
                for (uint32_t j = 0; j < L.ncand; j++) {
                    uint64_t p = L.place[c][j];
                    if (!p) continue;
                    uint64_t row = (p >> (8 * r)) & 0xffu;
                    if (!(row & ~(uint64_t)b)) m |= 1ull << j;
                }
                L.fit[((size_t)c * fit_rows + r) * 256 + b] = m;
            }
}

static void build_tables(d4_library &L, const uint64_t *shapes, const uint32_t *counts) {
    uint32_t k = 0, off = 0, g = 0, mins = UINT32_MAX;
    for (uint32_t i = 0; i < L.n; i++) {
        uint64_t s = shapes[i];
        if (!s) throw Error(D4_EINVAL, "empty piece shape");
        if (normalize8(s) != s) throw Error(D4_EINVAL, "piece shape not normalised");
        L.counts[i] = counts[i];
        L.sizes[i] = (uint32_t)__builtin_popcountll(s);
        L.fbits[i] = bitw(counts[i]);
        if ((off % 8) + L.fbits[i] > 8) off = (off + 7) & ~7u;  // no field straddles a byte
        L.off[i] = off;
        off += L.fbits[i];
        if (counts[i]) {
            uint32_t a = L.sizes[i], b = g;
            while (b) { uint32_t t = a % b; a = b; b = t; }
            g = a;
            mins = std::min(mins, L.sizes[i]);
        }
        // unique transforms enabled by sym, in order (Piece.cpp:12-22); the
        // transform id is the index within the sym-filtered list, as
        // Piece::cover passes it to make_step.
        std::vector<uint64_t> seen;
        uint32_t first = k;
        unsigned sym = L.sym, nall = 0;
        for (unsigned t = 0; t < 8; t++, sym >>= 1) {
            if (!(sym & 1u)) continue;
            uint64_t sh = transform8(s, t);
            bool dup = std::find(seen.begin(), seen.end(), sh) != seen.end();
            seen.push_back(sh);
            if (!dup) {
                if (k >= (uint32_t)MAX_CAND)
                    throw Error(D4_EINVAL, "more than 64 (piece, transform) pairs");
                L.cpiece[k] = (uint8_t)i;
                L.ctrs[k] = (uint8_t)nall;
                L.cnormal[k] = sh;
                L.cminc[k] = (uint8_t)__builtin_ctzll(sh & 0xffull);
                L.maxh = std::max(L.maxh, height8(sh));
                k++;
            }
            nall++;
        }
        uint64_t r = 0;
        for (uint32_t j = first; j < k; j++) r |= 1ull << j;
        for (uint32_t j = first; j < k; j++) L.range[j] = r;
        uint32_t info = L.off[i] | (((1u << L.fbits[i]) - 1u) << 8) | (counts[i] << 16);
        for (uint32_t j = first; j < k; j++) L.cinfo[j] = info;
        L.prange[i] = r;
    }
    if (off > 32) throw Error(D4_EINVAL, "piece counters need more than 32 bits");
    // xavail[b][v]: candidates of the pieces whose counter lies in byte b and
    // is exhausted when that byte reads v
    for (uint32_t b = 0; b < 4; b++)
        for (uint32_t v = 0; v < 256; v++) {
            uint64_t m = 0;
            for (uint32_t i = 0; i < L.n; i++) {
                if (L.off[i] / 8 != b && !(L.fbits[i] == 0 && b == 0)) continue;
                uint32_t f = L.fbits[i] ? (v >> (L.off[i] % 8)) & ((1u << L.fbits[i]) - 1u) : 0;
                if (f >= L.counts[i]) m |= L.prange[i];
            }
            L.xavail[b * 256 + v] = m;
        }
    if (L.maxh > (uint32_t)MAX_FIT_ROWS) throw Error(D4_EINVAL, "piece taller than 6 rows");
    L.used_bits = off;
    L.ncand = k;
    L.gcd = g ? g : 1;
    L.min_size = mins == UINT32_MAX ? 1 : mins;
    for (uint32_t c = 0; c < 8; c++)
        for (uint32_t j = 0; j < k; j++) {
            uint64_t n = L.cnormal[j];
            uint32_t mc = L.cminc[j];
            if (c >= mc && c - mc + width8(n) <= L.W) L.place[c][j] = n << (c - mc);
        }
    // packed per-anchor entries: shape, counter info and the piece's
    // candidate range [lo, hi)
    for (uint32_t c = 0; c < 8; c++)
        for (uint32_t j = 0; j < k; j++) {
            uint64_t r = L.range[j];
            uint64_t lo = (uint64_t)__builtin_ctzll(r), hi = 64 - (uint64_t)__builtin_clzll(r);
            L.placeinfo[(c * MAX_CAND + j) * 2] = L.place[c][j];
            L.placeinfo[(c * MAX_CAND + j) * 2 + 1] = (uint64_t)L.cinfo[j] | (lo << 32) | (hi << 40);
        }
    // the fit table covers the first fit_rows window rows (the placement
    // itself is tested in full when chosen); 3 rows keep it at 48 KiB of LDS
    build_fit(L, std::max(1u, std::min(L.maxh, 3u)));
    // comp[p]: ways to fill row pattern p with row-0 parts of placements,
    // ignoring lower rows and piece counts (bound used by next_size).
    L.comp[0] = 1;
    for (uint32_t p = 1; p < 256; p++) {
        uint32_t c = (uint32_t)__builtin_ctz(p);
        unsigned long long s = 0;
        for (uint32_t j = 0; j < k; j++) {
            uint64_t pl = L.place[c][j];
            if (!pl) continue;
            uint32_t r0 = (uint32_t)(pl & 0xffu);
            if (r0 & ~p) continue;
            s += L.comp[p & ~r0];
        }
        L.comp[p] = s;
    }
}

template <class T>
static void upload(T *&dst, const T *src, size_t n) {
    HIPCHK(hipMalloc((void **)&dst, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) HIPCHK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
}

// --------------------------------------------------------------------------
// frontier engine
// --------------------------------------------------------------------------

namespace {

struct RowLog {  // enum mode: what reconstruction needs from one row step
    std::vector<uint64_t> keys;       // frontier of this row (before the step)
    std::vector<uint32_t> e_parent;   // edges to the next row's frontier
    std::vector<uint64_t> e_desc;
    std::vector<uint32_t> e_child;
};

struct Engine {
    const d4_library *lib;
    uint32_t n_inst, mode, flags;
    std::vector<uint32_t> heights;
    std::vector<uint8_t> boards;  // [n_inst][64 + 8]
    uint32_t Hmax = 0, row = 0;
    bool done = false;
    // key layout
    uint32_t occ_bits = 0, inst_shift = 64;
    uint64_t occ_mask = 0;
    uint32_t used_mask = 0;
    uint32_t bw_stride = 0;
    // device state
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evx0 = nullptr, evx1 = nullptr, evc0 = nullptr;
    DevBuf<uint64_t> bw;
    DevBuf<uint32_t> hgt;
    DevBuf<unsigned long long> result, result_save;
    DevBuf<uint64_t> fk, fc, nk, nc;
    DevBuf<unsigned long long> tk;  // dedup table: {key, count} per slot
    DevBuf<uint32_t> c32;
    DevBuf<unsigned long long> c64;
    DevBuf<uint32_t> e_parent, e_child, slot_uid, blk;
    DevBuf<uint64_t> e_desc, e_slot;
    uint64_t n_front = 0;
    double grow_unique = 4.0, grow_children = 8.0;
    static constexpr uint64_t kExactCountMax = 1u << 16;
    d4_stats stats{};
    std::vector<unsigned long long> host_result;
    std::vector<RowLog> logs;
    std::vector<uint64_t> final_keys;

    Engine(const d4_library *l, uint32_t n, const uint8_t *b, uint32_t stride,
           const uint32_t *h, uint32_t m, uint32_t f)
        : lib(l), n_inst(n), mode(m), flags(f) {
        if (!n) throw Error(D4_EINVAL, "no instances");
        if (mode > D4_MODE_ENUM) throw Error(D4_EINVAL, "bad mode");
        heights.assign(h, h + n);
        boards.assign((size_t)n * (MAX_ROWS + 8), 0);
        for (uint32_t i = 0; i < n; i++) {
            if (heights[i] > (uint32_t)MAX_ROWS) throw Error(D4_EINVAL, "board taller than 64 rows");
            for (uint32_t r = 0; r < heights[i]; r++) {
                uint8_t v = b[(size_t)i * stride + r];
                if (lib->W < 8 && (v >> lib->W)) throw Error(D4_EINVAL, "board wider than the library's width");
                boards[(size_t)i * (MAX_ROWS + 8) + r] = v;
            }
            Hmax = std::max(Hmax, heights[i]);
        }
        occ_bits = 8 * (lib->maxh > 1 ? lib->maxh - 1 : 0);
        occ_mask = occ_bits ? (occ_bits >= 64 ? ~0ull : (1ull << occ_bits) - 1) : 0;
        used_mask = lib->used_bits >= 32 ? ~0u : (1u << lib->used_bits) - 1;
        uint32_t ib = n_inst > 1 ? bitw(n_inst - 1) : 0;
        if (occ_bits + lib->used_bits + ib > 63)
            throw Error(D4_EINVAL, "state does not fit a 63-bit key (occupancy " +
                                       std::to_string(occ_bits) + " + counters " +
                                       std::to_string(lib->used_bits) + " + instance " +
                                       std::to_string(ib) + " bits)");
        inst_shift = ib ? occ_bits + lib->used_bits : 64;
        HIPCHK(hipSetDevice(lib->device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        HIPCHK(hipEventCreate(&evx0));
        HIPCHK(hipEventCreate(&evx1));
        HIPCHK(hipEventCreate(&evc0));
        // board windows: bw[i][r] = rows r..r+7 of board i
        bw_stride = Hmax + 2;
        std::vector<uint64_t> hb((size_t)n * bw_stride, 0);
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t r = 0; r < bw_stride; r++) {
                uint64_t w = 0;
                for (uint32_t j = 0; j < 8 && r + j < (uint32_t)MAX_ROWS + 8; j++)
                    w |= (uint64_t)boards[(size_t)i * (MAX_ROWS + 8) + r + j] << (8 * j);
                hb[(size_t)i * bw_stride + r] = w;
            }
        bw.reserve(hb.size());
        HIPCHK(hipMemcpy(bw.p, hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
        hgt.reserve(n);
        HIPCHK(hipMemcpy(hgt.p, heights.data(), n * 4, hipMemcpyHostToDevice));
        result.reserve(n);
        HIPCHK(hipMemset(result.p, 0, n * 8));
        c32.reserve(C32_N);
        c64.reserve(C64_N);
        // root frontier: one empty-occupancy state per instance, count 1
        std::vector<uint64_t> k0(n), c0(n, 1);
        for (uint32_t i = 0; i < n; i++) k0[i] = inst_shift < 64 ? (uint64_t)i << inst_shift : 0;
        fk.reserve(n);
        fc.reserve(n);
        HIPCHK(hipMemcpy(fk.p, k0.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(fc.p, c0.data(), n * 8, hipMemcpyHostToDevice));
        n_front = n;
        stats.peak_frontier = n;
        if (Hmax == 0) row_step();
    }

    ~Engine() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : {evx0, evx1, evc0})
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    // rows still to be filled (CudaSearcher::get_height)
    uint32_t height() const { return done ? 0 : Hmax - row; }

    ExtendArgs args(uint64_t tab_m, uint32_t e_cap) {
        ExtendArgs a{};
        a.in_key = fk.p;
        a.in_cnt = fc.p;
        a.n_in = (uint32_t)n_front;
        // split levels: tiny frontiers and the heavy first rows spread each
        // record over 64 / 4096 work items
        a.split = n_front <= 1024 ? 2u : n_front <= (1u << 16) ? 1u : 0u;

        a.row = row;
        a.fit = lib->d_fit;
        a.placeinfo = lib->d_placeinfo;
        a.xavail = lib->d_xavail;
        a.fit_rows = lib->fit_rows;
        a.bw = bw.p;
        a.bw_stride = bw_stride;
        a.hgt = hgt.p;
        a.result = result.p;
        a.occ_bits = occ_bits;
        a.inst_shift = inst_shift;
        a.occ_mask = occ_mask;
        a.used_mask = used_mask;
        a.tab_key = tk.p;
        a.tab_cnt = tk.p + 1;  // interleaved {key, count} slots
        a.tab_mask = tab_m - 1;
        a.max_probe = 256;
        a.c32 = c32.p;
        a.c64 = c64.p;
        a.e_parent = e_parent.p;
        a.e_desc = e_desc.p;
        a.e_slot = e_slot.p;
        a.e_cap = e_cap;
        // rows (relative to the child's row) that no placed piece reaches
        a.seed_mask = lib->maxh >= 2 ? ~((1ull << (8 * (lib->maxh - 1))) - 1ull) : ~0ull;
        a.gcd = lib->gcd;
        a.min_size = lib->min_size;
        return a;
    }

    // One row step (search_GPU).  When the last row is done the terminal
    // states are absorbed into the per-instance counts in the same call.
    void step() {
        if (done) throw Error(D4_ESTATE, "search already finished");
        if (row < Hmax) row_step();
        if (!done && row >= Hmax) row_step();
    }

    // Rows 0..Hmax-1 extend the frontier; the step at row Hmax (and, for an
    // instance of height h, every step at row >= h) only absorbs terminal
    // states into the per-instance counts.
    void row_step() {
        if (n_front > 0xffffffffull) throw Error(D4_ERANGE, "frontier exceeds 2^32 records");
        const bool en = mode == D4_MODE_ENUM;
        const bool prune = !(flags & D4_FLAG_NO_PRUNE);
        // Table size: unique children <= children.  A small frontier gets an
        // exact count from a search-only pass; a large one is estimated from
        // the previous row's children per record (x2), with a rerun at 4x
        // the size if the table still overflows.
        uint64_t est_c = 0, est_u = 0;
        bool exact = false;
        if (n_front <= (1u << 20)) {
            // children <= sum over records of the row-0 completions (cheap)
            const uint64_t bound = next_size();
            if (bound <= (1ull << 26)) {
                est_c = est_u = bound;
                exact = true;
            } else if (n_front <= kExactCountMax) {
                est_c = est_u = count_children();
                stats.count_passes++;
                exact = true;
            }
        }
        if (!exact) {
            est_c = (uint64_t)(n_front * std::min(std::max(grow_children, 1.0), 64.0) * 1.5) + 4096;
            est_u = std::min<uint64_t>(est_c, (uint64_t)(n_front * std::max(grow_unique, 0.25) * 1.5) + 4096);
        }
        uint64_t tab_m = next_pow2(std::max<uint64_t>(2 * est_u, 1024));
        uint64_t est_e = en ? est_c + 64 : 0;
        if (en) {
            RowLog lg;
            lg.keys.resize(n_front);
            if (n_front)
                HIPCHK(hipMemcpy(lg.keys.data(), fk.p, n_front * 8, hipMemcpyDeviceToHost));
            logs.push_back(std::move(lg));
        }
        const bool absorbing = row >= *std::min_element(heights.begin(), heights.end());
        if (absorbing) result_save.reserve(n_inst);
        uint32_t h32[C32_N];
        unsigned long long h64[C64_N];
        for (;;) {
            if (tab_m > (1ull << 34)) throw Error(D4_ERANGE, "dedup table exceeds 2^34 slots");
            tk.reserve(2 * tab_m);
            if (en) {
                if (est_e > 0xffffffffull) throw Error(D4_ERANGE, "edge count exceeds 2^32");
                e_parent.reserve(est_e);
                e_desc.reserve(est_e);
                e_slot.reserve(est_e);
            }
            HIPCHK(hipEventRecord(ev0, st));
            if (absorbing)
                HIPCHK(hipMemcpyAsync(result_save.p, result.p, n_inst * 8,
                                      hipMemcpyDeviceToDevice, st));
            launch_fill_table(tk.p, tk.p + 1, tab_m, st);
            launch_reset_counters(c32.p, c64.p, st);
            HIPCHK(hipEventRecord(evx0, st));
            if (n_front)
                launch_extend(args(tab_m, (uint32_t)est_e), en, prune, false, lib->ext_threads,
                              lib->n_cu, st);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(evx1, st));
            HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (n_front) {
                float xms = 0;
                HIPCHK(hipEventElapsedTime(&xms, evx0, evx1));
                stats.extend_ms += xms;
                stats.extend_launches++;
                stats.extend_bytes += 16ull * n_front + 16ull * h64[C64_CHILDREN];
            }
            bool ovf = h32[C32_OVERFLOW] != 0;
            bool eovf = en && h32[C32_EDGE_OVF] != 0;
            if (!ovf && !eovf) break;
            stats.retries++;
            // the step is a pure function of the frontier: rerun it bigger,
            // after undoing the terminal counts it already added
            if (absorbing)
                HIPCHK(hipMemcpy(result.p, result_save.p, n_inst * 8,
                                 hipMemcpyDeviceToDevice));
            if (ovf) tab_m *= 4;
            if (eovf) est_e = (uint64_t)h32[C32_EDGES] + 1024;
        }
        const uint64_t U = h64[C64_NEW];
        stats.children += h64[C64_CHILDREN];
        stats.pruned += h64[C64_PRUNED];
        stats.records_extended += n_front;
        stats.unique += U;
        if (n_front) {
            grow_unique = std::max(1.0, (double)U / (double)n_front);
            grow_children = std::max(1.0, (double)h64[C64_CHILDREN] / (double)n_front);
        }
        nk.reserve(U);
        nc.reserve(U);
        if (en) slot_uid.reserve(tab_m);
        HIPCHK(hipEventRecord(evc0, st));
        blk.reserve(2 * COMPACT_MAX_BLOCKS);
        launch_compact(tk.p, tk.p + 1, tab_m, nk.p, nc.p, c32.p, blk.p, en ? slot_uid.p : nullptr,
                       st);
        uint32_t n_edges = en ? h32[C32_EDGES] : 0;
        if (en && n_edges) {
            e_child.reserve(n_edges);
            launch_remap_edges(e_slot.p, slot_uid.p, e_child.p, n_edges, st);
        }
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        float ms = 0, cms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&cms, evc0, ev1));
        stats.gpu_ms += ms;
        stats.compact_ms += cms;
        stats.compact_launches++;
        stats.compact_bytes += 8ull * tab_m + 24ull * U;
        if (en) {
            RowLog &lg = logs.back();
            lg.e_parent.resize(n_edges);
            lg.e_desc.resize(n_edges);
            lg.e_child.resize(n_edges);
            if (n_edges) {
                HIPCHK(hipMemcpy(lg.e_parent.data(), e_parent.p, n_edges * 4, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(lg.e_desc.data(), e_desc.p, n_edges * 8, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(lg.e_child.data(), e_child.p, n_edges * 4, hipMemcpyDeviceToHost));
            }
        }
        fk.swap(nk);
        fc.swap(nc);
        n_front = U;
        if (!U) fk.reserve(1);
        stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, U);
        stats.rows_done++;
        if (row >= Hmax) {
            done = true;
            host_result.resize(n_inst);
            HIPCHK(hipMemcpy(host_result.data(), result.p, n_inst * 8, hipMemcpyDeviceToHost));
        } else {
            row++;
        }
    }

    void run() {
        while (!done) step();
    }

    // Children the current frontier will emit (after pruning): the extend
    // search with inserts switched off.
    uint64_t count_children() {
        if (!n_front) return 0;
        const bool prune = !(flags & D4_FLAG_NO_PRUNE);
        launch_reset_counters(c32.p, c64.p, st);
        launch_extend(args(1, 0), false, prune, true, lib->ext_threads, lib->n_cu, st);
        HIPCHK(hipGetLastError());
        unsigned long long h64[C64_N];
        HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return h64[C64_CHILDREN];
    }

    uint64_t next_size() {
        if (done || !n_front) return 0;
        launch_reset_counters(c32.p, c64.p, st);
        launch_next_size(fk.p, (uint32_t)n_front, row, bw.p, bw_stride, hgt.p, inst_shift,
                         occ_mask, lib->d_comp, c32.p, c64.p, lib->n_cu, st);
        unsigned long long h64[C64_N];
        HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return h64[C64_NEXTSIZE];
    }

    // ---- reconstruction (enum mode) --------------------------------------

    // Each solution as its sequence of candidate ids in placement order.
    // Walks the back-pointers from instance i's terminal states (the
    // frontier at row heights[i]) down to the root.
    std::vector<std::vector<uint8_t>> solutions_of(uint32_t inst, bool single) const {
        if (mode != D4_MODE_ENUM) throw Error(D4_ESTATE, "solutions need D4_MODE_ENUM");
        if (!done) throw Error(D4_ESTATE, "search not finished");
        const uint32_t H = heights[inst];
        std::vector<std::vector<uint8_t>> out;
        auto inst_of = [&](uint64_t key) -> uint32_t {
            return inst_shift < 64 ? (uint32_t)(key >> inst_shift) : 0u;
        };
        // reverse adjacency per row: child -> edges
        std::vector<std::vector<uint32_t>> start(H), order(H);
        for (uint32_t r = 0; r < H; r++) {
            const RowLog &lg = logs[r];
            size_t nchild = r + 1 < logs.size() ? logs[r + 1].keys.size() : 0;
            std::vector<uint32_t> cnt(nchild + 1, 0);
            for (uint32_t c : lg.e_child) cnt[c + 1]++;
            for (size_t j = 1; j <= nchild; j++) cnt[j] += cnt[j - 1];
            std::vector<uint32_t> ord(lg.e_child.size());
            std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
            for (uint32_t e = 0; e < lg.e_child.size(); e++) ord[pos[lg.e_child[e]]++] = e;
            start[r] = std::move(cnt);
            order[r] = std::move(ord);
        }
        std::vector<uint64_t> path(H);
        std::vector<std::vector<uint64_t>> paths;
        std::function<void(uint32_t, uint32_t)> walk = [&](uint32_t r, uint32_t u) {
            // u indexes the frontier of row r (r >= 1); edges of row r-1 lead here
            if (r == 0) {
                paths.push_back(path);
                return;
            }
            const RowLog &lg = logs[r - 1];
            for (uint32_t j = start[r - 1][u]; j < start[r - 1][u + 1]; j++) {
                uint32_t e = order[r - 1][j];
                path[r - 1] = lg.e_desc[e];
                walk(r - 1, lg.e_parent[e]);
            }
        };
        if (H == 0) {
            paths.push_back({});
        } else if (H < logs.size()) {
            const std::vector<uint64_t> &term = logs[H].keys;
            for (uint32_t u = 0; u < term.size(); u++)
                if (inst_of(term[u]) == inst) walk(H, u);
        }
        for (auto &p : paths) {
            std::vector<uint8_t> seq;
            for (uint64_t d : p)
                for (int j = 0; j < 8 && ((d >> (8 * j)) & 0xffu); j++)
                    seq.push_back((uint8_t)(((d >> (8 * j)) & 0xffu) - 1));
            out.push_back(std::move(seq));
        }
        // the CPU searcher's DFS visits candidates in id order at every
        // anchor, so its order is the lexicographic order of these sequences
        std::sort(out.begin(), out.end());
        if (single && out.size() > 1) out.resize(1);
        return out;
    }

    // Steps {piece_id, trs_id, y, x} of a candidate sequence, replayed on the
    // board (anchor = lowest open cell, Piece.inl:5-24).
    void encode(uint32_t inst, const std::vector<uint8_t> &seq, std::vector<uint8_t> &buf) const {
        uint8_t open[MAX_ROWS + 8];
        std::memcpy(open, &boards[(size_t)inst * (MAX_ROWS + 8)], sizeof open);
        buf.push_back((uint8_t)seq.size());
        uint32_t r = 0;
        for (uint8_t k : seq) {
            while (r < heights[inst] && !open[r]) r++;
            if (r >= heights[inst]) throw Error(D4_EDEVICE, "inconsistent back-pointers");
            uint32_t c = (uint32_t)__builtin_ctz(open[r]);
            uint64_t pl = lib->place[c][k];
            for (uint32_t j = 0; j < 8 && r + j < (uint32_t)MAX_ROWS + 8; j++) {
                uint8_t rowbits = (uint8_t)(pl >> (8 * j));
                if (rowbits & ~open[r + j]) throw Error(D4_EDEVICE, "inconsistent back-pointers");
                open[r + j] &= (uint8_t)~rowbits;
            }
            buf.push_back(lib->cpiece[k]);
            buf.push_back(lib->ctrs[k]);
            buf.push_back((uint8_t)r);
            buf.push_back((uint8_t)(c - lib->cminc[k]));
        }
    }

    void solutions(uint32_t inst, int single, uint8_t *out, size_t cap, size_t *len,
                   uint64_t *nsol) const {
        if (inst >= n_inst) throw Error(D4_EINVAL, "instance out of range");
        auto seqs = solutions_of(inst, single != 0);
        std::vector<uint8_t> buf;
        for (auto &s : seqs) encode(inst, s, buf);
        if (len) *len = buf.size();
        if (nsol) *nsol = seqs.size();
        if (out && buf.size() <= cap) std::memcpy(out, buf.data(), buf.size());
        else if (out && cap) throw Error(D4_ERANGE, "solution buffer too small");
    }
};

}  // namespace

struct d4_searcher {
    std::unique_ptr<Engine> e;
};
struct d4_batch {
    std::unique_ptr<Engine> e;
};

// --------------------------------------------------------------------------
// C ABI
// --------------------------------------------------------------------------

extern "C" {

const char *d4_last_error(void) { return g_err.c_str(); }

int d4_device_count(int *n) {
    return guard([&] {
        if (!n) throw Error(D4_EINVAL, "null argument");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *n = e == hipSuccess ? c : 0;
    });
}

int d4_device_synchronize(void) {
    return guard([&] { HIPCHK(hipDeviceSynchronize()); });
}

size_t d4_free_polyominoes(unsigned n, uint64_t *out, size_t cap) {
    if (n < 1 || n > 8) return 0;
    // grow every (n-1)-cell free polyomino by one 4-neighbour cell and keep
    // the canonical forms
    std::vector<uint64_t> cur{1};
    for (unsigned k = 2; k <= n; k++) {
        std::vector<uint64_t> nxt;
        for (uint64_t s0 : cur) {
            uint64_t s = s0 << 9;  // one free row above and column left
            for (uint64_t b = s; b; b &= b - 1) {
                int i = __builtin_ctzll(b), r = i / 8, c = i % 8;
                const int nb[4][2] = {{r - 1, c}, {r + 1, c}, {r, c - 1}, {r, c + 1}};
                for (auto &q : nb) {
                    if (q[0] < 0 || q[0] > 7 || q[1] < 0 || q[1] > 7) continue;
                    uint64_t bit = 1ull << (q[0] * 8 + q[1]);
                    if (s & bit) continue;
                    uint64_t v = ~0ull;
                    for (unsigned t = 0; t < 8; t++) v = std::min(v, transform8(s | bit, t));
                    nxt.push_back(v);
                }
            }
        }
        std::sort(nxt.begin(), nxt.end());
        nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
        cur.swap(nxt);
    }
    for (size_t i = 0; i < cur.size() && i < cap && out; i++) out[i] = cur[i];
    return cur.size();
}

int d4_library_create(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                      uint32_t sym, uint32_t width, d4_library **out) {
    return guard([&] {
        if (!out || !shapes || !counts) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        if (!n || n > (uint32_t)MAX_PIECES) throw Error(D4_EINVAL, "need 1..32 pieces");
        if (!width || width > 8) throw Error(D4_EINVAL, "width must be 1..8");
        if (!(sym & 0xffu)) throw Error(D4_EINVAL, "empty symmetry mask");
        auto L = std::make_unique<d4_library>();
        L->n = n;
        L->sym = sym & 0xffu;
        L->W = width;
        build_tables(*L, shapes, counts);
        int dev = 0, ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (!ndev) throw Error(D4_EDEVICE, "no HIP device");
        HIPCHK(hipGetDevice(&dev));
        L->device = dev;
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        L->n_cu = prop.multiProcessorCount;
        L->lds_limit = (uint32_t)prop.sharedMemPerBlock;
        if (extend_lds_bytes(L->fit_rows, 1024, true) > L->lds_limit)
            throw Error(D4_EINVAL, "placement tables exceed LDS");
        upload(L->d_fit, L->fit.data(), L->fit.size());
        upload(L->d_placeinfo, L->placeinfo, 2 * 8 * MAX_CAND);
        upload(L->d_xavail, L->xavail, 4 * 256);
        upload(L->d_comp, L->comp, 256);
        *out = L.release();
    });
}

void d4_library_destroy(d4_library *lib) { delete lib; }

int d4_library_info(const d4_library *lib, uint32_t *n_cand, uint32_t *max_height) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (n_cand) *n_cand = lib->ncand;
        if (max_height) *max_height = lib->maxh;
    });
}

int d4_library_set_launch(d4_library *lib, uint32_t fit_rows, uint32_t threads) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (threads != 256 && threads != 512 && threads != 1024)
            throw Error(D4_EINVAL, "threads must be 256, 512 or 1024");
        if (fit_rows < 1 || fit_rows > lib->maxh) throw Error(D4_EINVAL, "fit_rows out of range");
        if (extend_lds_bytes(fit_rows, threads, true) > lib->lds_limit)
            throw Error(D4_EINVAL, "launch shape exceeds LDS");
        if (fit_rows != lib->fit_rows) {
            HIPCHK(hipSetDevice(lib->device));
            build_fit(*lib, fit_rows);
            uint64_t *old = lib->d_fit;
            lib->d_fit = nullptr;
            upload(lib->d_fit, lib->fit.data(), lib->fit.size());
            if (old) (void)hipFree(old);
        }
        lib->ext_threads = threads;
    });
}

int d4_library_get_launch(const d4_library *lib, uint32_t *fit_rows, uint32_t *threads) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (fit_rows) *fit_rows = lib->fit_rows;
        if (threads) *threads = lib->ext_threads;
    });
}

int d4_library_tune(d4_library *lib, uint32_t n_inst, const uint8_t *boards, uint32_t stride,
                    const uint32_t *heights, uint32_t *fit_rows, uint32_t *threads,
                    double *best_ms) {
    return guard([&] {
        if (!lib || !boards || !heights || !n_inst) throw Error(D4_EINVAL, "null argument");
        uint32_t best_f = lib->fit_rows, best_t = lib->ext_threads;
        double best = -1;
        {  // warm the allocator cache and the code objects
            Engine e(lib, n_inst, boards, stride, heights, D4_MODE_COUNT, 0);
            e.run();
        }
        for (uint32_t f = 1; f <= std::min(lib->maxh, 4u); f++)
            for (uint32_t t : {256u, 512u, 1024u}) {
                if (extend_lds_bytes(f, t, true) > lib->lds_limit) continue;
                if (d4_library_set_launch(lib, f, t) != D4_OK) continue;
                Engine e(lib, n_inst, boards, stride, heights, D4_MODE_COUNT, 0);
                e.run();
                double ms = e.stats.extend_ms;
                if (best < 0 || ms < best) {
                    best = ms;
                    best_f = f;
                    best_t = t;
                }
            }
        if (d4_library_set_launch(lib, best_f, best_t) != D4_OK)
            throw Error(D4_EDEVICE, g_err);
        if (fit_rows) *fit_rows = best_f;
        if (threads) *threads = best_t;
        if (best_ms) *best_ms = best;
    });
}

int d4_searcher_create(const d4_library *lib, const uint8_t *rows, uint32_t height,
                       uint32_t mode, uint32_t flags, d4_searcher **out) {
    return guard([&] {
        if (!lib || !out || (!rows && height)) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        auto s = std::make_unique<d4_searcher>();
        s->e = std::make_unique<Engine>(lib, 1, rows, height, &height, mode, flags);
        *out = s.release();
    });
}

int d4_searcher_create_ea(const d4_library *lib, uint64_t empty_area, uint32_t mode,
                          uint32_t flags, d4_searcher **out) {
    uint8_t rows[8];
    uint32_t bw = empty_area ? 64u - (uint32_t)__builtin_clzll(empty_area) : 0u;
    uint32_t h = (bw + 7) / 8;  // CudaSearcher: height = ceil(bit_width / 8)
    for (int r = 0; r < 8; r++) rows[r] = (uint8_t)(empty_area >> (8 * r));
    return d4_searcher_create(lib, rows, h, mode, flags, out);
}

void d4_searcher_destroy(d4_searcher *s) { delete s; }

uint32_t d4_searcher_height(const d4_searcher *s) { return s ? s->e->height() : 0; }

uint64_t d4_searcher_size(const d4_searcher *s) { return s ? s->e->n_front : 0; }

uint64_t d4_searcher_next_size(const d4_searcher *s) {
    uint64_t v = 0;
    if (!s) return 0;
    if (guard([&] { v = s->e->next_size(); }) != D4_OK) return 0;
    return v;
}

int d4_searcher_step(d4_searcher *s) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        s->e->step();
    });
}

int d4_searcher_run(d4_searcher *s) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        s->e->run();
    });
}

int d4_searcher_count(const d4_searcher *s, uint64_t *count) {
    return guard([&] {
        if (!s || !count) throw Error(D4_EINVAL, "null argument");
        if (!s->e->done) throw Error(D4_ESTATE, "search not finished");
        *count = s->e->host_result[0];
    });
}

int d4_searcher_solutions(const d4_searcher *s, int single, uint8_t *out, size_t cap,
                          size_t *len, uint64_t *nsol) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        s->e->solutions(0, single, out, cap, len, nsol);
    });
}

int d4_searcher_stats(const d4_searcher *s, d4_stats *st) {
    return guard([&] {
        if (!s || !st) throw Error(D4_EINVAL, "null argument");
        *st = s->e->stats;
    });
}

int d4_batch_create(const d4_library *lib, uint32_t n_inst, const uint8_t *boards,
                    uint32_t stride, const uint32_t *heights, uint32_t mode, uint32_t flags,
                    d4_batch **out) {
    return guard([&] {
        if (!lib || !out || !boards || !heights) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        auto b = std::make_unique<d4_batch>();
        b->e = std::make_unique<Engine>(lib, n_inst, boards, stride, heights, mode, flags);
        *out = b.release();
    });
}

void d4_batch_destroy(d4_batch *b) { delete b; }

int d4_batch_run(d4_batch *b) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        b->e->run();
    });
}

int d4_batch_counts(const d4_batch *b, uint64_t *counts) {
    return guard([&] {
        if (!b || !counts) throw Error(D4_EINVAL, "null argument");
        if (!b->e->done) throw Error(D4_ESTATE, "batch not finished");
        for (uint32_t i = 0; i < b->e->n_inst; i++) counts[i] = b->e->host_result[i];
    });
}

int d4_batch_stats(const d4_batch *b, d4_stats *st) {
    return guard([&] {
        if (!b || !st) throw Error(D4_EINVAL, "null argument");
        *st = b->e->stats;
    });
}

int d4_batch_solutions(const d4_batch *b, uint32_t i, int single, uint8_t *out, size_t cap,
                       size_t *len, uint64_t *nsol) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        b->e->solutions(i, single, out, cap, len, nsol);
    });
}

}  // extern "C"
