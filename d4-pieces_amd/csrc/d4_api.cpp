// d4_api.cpp -- host side of the row-by-row tiling search: placement tables,
// the frontier engine that drives the gfx950 kernels one row at a time, the
// reconstruction of solutions from back-pointers, and the C ABI (d4.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "d4_host.h"
#include "d4_library.h"

namespace d4 {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

}  // namespace d4

using namespace d4;

// --------------------------------------------------------------------------
// piece library and placement tables
// --------------------------------------------------------------------------


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

// Piece counters, candidates and the pruning constants (any library of up
// to MAX_LIB_CAND candidates).
static void build_tables(d4_library &L, const uint64_t *shapes, const uint32_t *counts,
                         const uint8_t *enabled) {
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
        // Piece::cover passes it to make_step.  An `enabled` mask switches
        // placements on by that index instead (Piece::Placement::enabled,
        // duplicates included when their bit is set).
        std::vector<uint64_t> seen;
        uint32_t first = k;
        unsigned sym = L.sym, nall = 0;
        for (unsigned t = 0; t < 8; t++, sym >>= 1) {
            if (!(sym & 1u)) continue;
            uint64_t sh = transform8(s, t);
            bool dup = std::find(seen.begin(), seen.end(), sh) != seen.end();
            seen.push_back(sh);
            if (enabled ? ((enabled[i] >> nall) & 1u) != 0 : !dup) {
                if (k >= (uint32_t)MAX_LIB_CAND)
                    throw Error(D4_EINVAL, "more than 128 (piece, transform) pairs");
                L.cpiece[k] = (uint8_t)i;
                L.ctrs[k] = (uint8_t)nall;
                L.cnormal[k] = sh;
                L.cminc[k] = (uint8_t)__builtin_ctzll(sh & 0xffull);
                L.maxh = std::max(L.maxh, height8(sh));
                k++;
            }
            nall++;
        }
        uint32_t info = L.off[i] | (((1u << L.fbits[i]) - 1u) << 8) | (counts[i] << 16);
        for (uint32_t j = first; j < k; j++) {
            L.cinfo[j] = info;
            L.clo[j] = (uint8_t)first;
            L.chi[j] = (uint8_t)k;
        }
    }
    if (off > 32) throw Error(D4_EINVAL, "piece counters need more than 32 bits");
    if (L.maxh > (uint32_t)MAX_FIT_ROWS) throw Error(D4_EINVAL, "piece taller than 6 rows");
    L.used_bits = off;
    L.ncand = k;
    L.gcd = g ? g : 1;
    L.min_size = mins == UINT32_MAX ? 1 : mins;
    // the fit table covers the first fit_rows window rows (the placement
    // itself is tested in full when chosen); 3 rows keep it at 48 KiB of LDS
    L.fit_rows = std::max(1u, std::min(L.maxh, 3u));
}

// Single-library tables of the LDS-staged extend kernel: u64 candidate masks
// per anchor column and window row, counter-byte availability tables, packed
// placement entries and the row-0 completion counts used by next_size.
static void build_fast(d4_library &L) {
    const uint32_t k = L.ncand;
    for (uint32_t j = 0; j < k; j++) {
        uint64_t r = 0;
        for (uint32_t q = L.clo[j]; q < L.chi[j]; q++) r |= 1ull << q;
        L.range[j] = r;
        L.prange[L.cpiece[j]] = r;
    }
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
            L.placeinfo[(c * MAX_CAND + j) * 2] = L.place[c][j];
            L.placeinfo[(c * MAX_CAND + j) * 2 + 1] =
                (uint64_t)L.cinfo[j] | ((uint64_t)L.clo[j] << 32) | ((uint64_t)L.chi[j] << 40);
        }
    build_fit(L, L.fit_rows);
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
    if (n) {
        // complete before any engine's (non-blocking) stream reads the table
        HIPCHK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipStreamSynchronize(nullptr));
    }
}

// Build and upload the single-library tables once (engines call this).
static void ensure_fast(const d4_library *cl) {
    d4_library *L = const_cast<d4_library *>(cl);
    std::lock_guard<std::mutex> lk(L->fast_mu);
    if (L->fast_ready) return;
    if (L->ncand > (uint32_t)MAX_CAND)
        throw Error(D4_EINVAL, "more than 64 (piece, transform) pairs: this library runs "
                               "only in a multi-library batch (d4_mbatch_create)");
    build_fast(*L);
    if (extend_lds_bytes(L->fit_rows, 256, true) > L->lds_limit)
        throw Error(D4_EINVAL, "placement tables exceed LDS");
    HIPCHK(hipSetDevice(L->device));
    upload(L->d_fit, L->fit.data(), L->fit.size());
    upload(L->d_placeinfo, L->placeinfo, 2 * 8 * MAX_CAND);
    upload(L->d_xavail, L->xavail, 4 * 256);
    upload(L->d_comp, L->comp, 256);
    L->fast_ready = true;
}

// --------------------------------------------------------------------------
// allocation-free scratch (d4_searcher_count_noalloc)
// --------------------------------------------------------------------------

// The allocation-free entry point's device memory: one array per device for
// the whole process, carved into the count-mode engine's buffers.  Host code
// must never pass `scratch` itself to a kernel or copy (in host code the name
// is a stand-in, not a device address): the library looks up each device's
// address once, at library creation, with hipGetSymbolAddress.
constexpr uint64_t kScratchBytes = 128ull << 20;
__device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
// slices of the scratch, in entries
constexpr uint64_t kNaFront = 512u << 10;  // frontier states (x4 arrays: current, next)
constexpr uint64_t kNaTable = 2u << 20;    // dedup table slots (16 B)
constexpr uint64_t kNaLevel = 512u << 10;  // partial rows per level buffer (x2, 40 B)
constexpr uint64_t kNaChild = 2u << 20;    // children per chunk (16 B)
constexpr uint64_t kNaSpill = 16u << 10;   // spilled children (16 B)

struct NoAllocDev {
    std::mutex mu;  // serialises this process's calls on the device
    bool ready = false;
    uint8_t *ptr = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[5] = {};
};
NoAllocDev g_noalloc[kMaxDevices];
std::mutex g_noalloc_mu;

// Once per device, with that device current: the scratch's address on it,
// and the stream and events the entry point runs on.
void noalloc_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoAllocDev &d = g_noalloc[dev];
    if (d.ready) return;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HIPCHK(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking));
    for (hipEvent_t &e : d.ev) HIPCHK(hipEventCreate(&e));
    d.ptr = static_cast<uint8_t *>(p);
    d.ready = true;
}

}  // namespace

// --------------------------------------------------------------------------
// frontier engine
// --------------------------------------------------------------------------

namespace {

struct RowLog {  // enum mode: one global step, copied to the host
    std::vector<uint64_t> keys;       // frontier before the step
    std::vector<uint32_t> e_parent;   // edges to the next step's frontier
    std::vector<uint64_t> e_desc;
    std::vector<uint32_t> e_child;
};

struct DevLog {  // enum mode: one global step, kept on the device
    DevBuf<uint64_t> keys;  // frontier before the step
    uint64_t n_keys = 0;
    DevBuf<uint32_t> par, chi;  // edges to the next step's frontier
    DevBuf<uint64_t> desc;
    uint64_t n_edges = 0;
};

// The frontier engine: S instance slots sharing one library and one device
// frontier.  An instance joins at a global step (its "start") and advances
// one of its rows per step; it finishes at step start + height, when its
// terminal states are absorbed into its count.  A batch admits every
// instance at step 0; a server admits requests between steps (continuous
// batching).
struct Engine {
    const d4_library *lib;
    uint32_t S, mode, flags, max_h;
    bool keep_logs;
    // per slot (host side)
    std::vector<uint32_t> heights, start;
    std::vector<uint8_t> boards;  // [S][MAX_ROWS + 8]
    std::vector<uint8_t> active;
    std::vector<unsigned long long> host_result;
    std::vector<uint32_t> pending, completed;
    uint32_t g = 0, n_active = 0;
    // key layout
    uint32_t occ_bits = 0, inst_shift = 64;
    uint64_t occ_mask = 0;
    uint32_t used_mask = 0;
    uint32_t bw_stride = 0;
    // device state
    hipStream_t st = nullptr;

    // Every copy of an engine runs on its own stream: the stream is
    // non-blocking, so a copy on the null stream would not be ordered with
    // the engine's kernels (a count restore could land after the rerun it
    // prepares).  The host waits for it, so host buffers may be reused.
    // Waits for the engine's stream (counted: host round trips per row step).
    void sync_host() const {
        HIPCHK(hipStreamSynchronize(st));
        stats.host_syncs++;
    }
    void copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) const {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
        sync_host();
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evx0 = nullptr, evx1 = nullptr, evc0 = nullptr;
    DevBuf<uint64_t> bw;
    DevBuf<uint32_t> hgt, dstart;
    DevBuf<unsigned long long> result, result_save;
    DevBuf<uint64_t> fk, fc, nk, nc;
    DevBuf<unsigned long long> tk;  // dedup table: {key, count} per slot
    DevBuf<uint32_t> c32;
    DevBuf<unsigned long long> c64;
    DevBuf<uint32_t> e_parent, e_child, slot_uid, blk;
    DevBuf<uint64_t> e_desc, e_slot;
    DevBuf<uint64_t> don_a, don_b;  // donated work items (two words each)
    // k_expand's spilled partial-row states (EXPAND_WORDS words each), items per buffer
    uint64_t xp_cap = 0;  // (set from the library's donation capacity)
    // level kernels: partial-row buffers (two, alternating), the children
    // buffer, and per-parent sizes seen on the last row step
    DevBuf<uint64_t> lv_a, lv_b, ch_k, ch_c;
    // partitioned dedup (level kernels, count mode): children in bucket
    // order, per (bucket, tile) counts / offsets, scan block sums
    DevBuf<uint64_t> pbuf, sbuf, pmask;
    DevBuf<uint32_t> phist, phist2, pbsum;
    bool last_part = false;  // the last fused step deduped by partitions
    double lv_ratio = 0.0, ch_ratio = 0.0;  // max level / children written per parent
    double pp_bound = 0.0;                   // children bound per parent (this step)
    double seen_lv = 0.0, seen_ch = 0.0;     // the library's sizes for this step index
    uint64_t lv_cap = 0, ch_cap = 0;  // last sizes used
    std::vector<hipEvent_t> lv_ev;  // chunk boundaries: levels / dedup timing
    // fused count-mode row step: the whole chain queued, one host round trip
    uint32_t *pin32 = nullptr;              // pinned copies of the counters
    unsigned long long *pin64 = nullptr;
    static constexpr uint64_t kFusedMaxSlots = 1ull << 30;
    // graph-replayed row steps: two recorded steps (frontier buffer 0 -> 1
    // and 1 -> 0) whose kernels read their sizes from a device control block
    hipGraphExec_t gx[2] = {nullptr, nullptr};
    DevBuf<uint64_t> gk[2], gc[2];       // the frontier buffers they alternate between
    DevBuf<RowCtl> ctl_d;
    RowCtl *ctl_h = nullptr;             // pinned
    std::vector<uint64_t> g_sig;         // buffers and sizes the graphs were recorded with
    uint64_t g_cap = 0, g_lv = 0, g_ch = 0;
    bool g_part = false;  // the recorded steps dedup by partitions
    uint32_t g_fail = 0;                 // the last replay's failure flags
    bool g_off_once = false;             // the next row step does not replay
    static constexpr uint64_t kGraphMaxSlots = 1ull << 25;
    uint64_t n_front = 0, new_roots = 0;
    double grow_unique = 4.0, grow_children = 8.0, grow_unique_raw = 4.0;
    double prev_unique_raw = 0.0;  // the step before's unique ratio (0: none yet)
    DevBuf<unsigned long long> tk2;  // the larger table a spill moves into
    DevBuf<uint64_t> spill;          // count mode: children that found no slot
    static constexpr uint64_t kSpillCap = 1u << 22;
    // bounded memory (count mode, run()): frontier halves parked for later
    uint64_t mem_cap = 0;  // 0: unbounded
    // A parked part lies in device memory, or, with a host budget set, in
    // pinned host memory (keys then counts), so that the frontier of a row
    // can exceed the device: halves stream out over the host link and back
    // when their turn comes.
    struct Part {
        uint32_t g = 0;
        uint64_t n = 0;
        double grow_unique = 1.0, grow_children = 1.0, grow_unique_raw = 1.0;
        double prev_unique_raw = 0.0;
        DevBuf<uint64_t> k, c;
        uint64_t *host = nullptr;  // [2][n] when parked on the host
        ~Part() {
            if (host) (void)hipHostFree(host);
        }
    };
    std::vector<std::unique_ptr<Part>> parked;
    uint64_t parked_bytes = 0;       // device bytes of the parked parts
    uint64_t host_cap = 0;           // pinned host bytes parts may take (0: none)
    uint64_t host_parked = 0;
    bool stream_ok = true;           // rows that do not fit stream (else halves are parked)
    static constexpr uint32_t kMaxPartitions = 64;
    // The frontier in parts (device or pinned host memory) as a streamed row
    // step left it when it was too large to gather into fk / fc; empty:
    // fk / fc hold the frontier.
    std::vector<std::unique_ptr<Part>> fparts;
    DevBuf<uint64_t> sk, sc;         // a host part's chunk, copied in
    bool trace = getenv("D4_TRACE") != nullptr;  // per-step lines of run() on stderr
    // scale of the count-mode dedup table estimate (D4_TABLE_SCALE, default
    // 1): below 1 the tables overflow and spill, which exercises the reruns
    double table_scale = [] {
        const char *v = getenv("D4_TABLE_SCALE");
        const double x = v ? atof(v) : 1.0;
        return x > 0 ? std::min(x, 16.0) : 1.0;
    }();
    static constexpr uint64_t kExactCountMax = 1u << 16;
    // children per record allowed for when a first row is sized without a
    // count (count mode; more spill into a grown table)
    static constexpr uint64_t kFirstRowFanout = 8192;
    mutable d4_stats stats{};  // mutable: const reads of device logs count their waits
    std::deque<DevLog> logs;  // logs[j] = global step log0 + j
    uint32_t log0 = 0;

    Engine(const d4_library *l, uint32_t slots, uint32_t max_height, uint32_t m, uint32_t f,
           bool keep)
        : lib(l), S(slots), mode(m), flags(f), max_h(max_height), keep_logs(keep) {
        if (!S) throw Error(D4_EINVAL, "no instances");
        if (mode > D4_MODE_ENUM) throw Error(D4_EINVAL, "bad mode");
        ensure_fast(lib);
        if (max_h > (uint32_t)MAX_ROWS) throw Error(D4_EINVAL, "board taller than 64 rows");
        heights.assign(S, 0);
        start.assign(S, 0);
        active.assign(S, 0);
        host_result.assign(S, 0);
        boards.assign((size_t)S * (MAX_ROWS + 8), 0);
        occ_bits = 8 * (lib->maxh > 1 ? lib->maxh - 1 : 0);
        occ_mask = occ_bits ? (occ_bits >= 64 ? ~0ull : (1ull << occ_bits) - 1) : 0;
        used_mask = lib->used_bits >= 32 ? ~0u : (1u << lib->used_bits) - 1;
        uint32_t ib = S > 1 ? bitw(S - 1) : 0;
        if (occ_bits + lib->used_bits + ib > 63)
            throw Error(D4_EINVAL, "state does not fit a 63-bit key (occupancy " +
                                       std::to_string(occ_bits) + " + counters " +
                                       std::to_string(lib->used_bits) + " + instance " +
                                       std::to_string(ib) + " bits)");
        inst_shift = ib ? occ_bits + lib->used_bits : 64;
        xp_cap = lib->don_items;
        HIPCHK(hipSetDevice(lib->device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        for (hipEvent_t *e : {&ev0, &ev1, &evx0, &evx1, &evc0}) HIPCHK(hipEventCreate(e));
        bw_stride = max_h + 2;
        bw.reserve((uint64_t)S * bw_stride);
        hgt.reserve(S);
        dstart.reserve(S);
        result.reserve(S);
        HIPCHK(hipMemsetAsync(bw.p, 0, (uint64_t)S * bw_stride * 8, st));
        HIPCHK(hipMemsetAsync(hgt.p, 0, (uint64_t)S * 4, st));
        HIPCHK(hipMemsetAsync(dstart.p, 0, (uint64_t)S * 4, st));
        HIPCHK(hipMemsetAsync(result.p, 0, (uint64_t)S * 8, st));
        c32.reserve(C32_N);
        c64.reserve(C64_N);
        fk.reserve(S);
        fc.reserve(S);
        HIPCHK(hipStreamSynchronize(st));
    }

    // A count-mode engine for one board whose buffers are slices of the
    // device's allocation-free scratch, on the stream and events made for
    // it at set-up: nothing is allocated on the device.
    bool noalloc = false;
    Engine(const d4_library *l, uint32_t max_height, uint32_t f, NoAllocDev &na)
        : lib(l), S(1), mode(D4_MODE_COUNT), flags(f), max_h(max_height), keep_logs(false) {
        if (!lib->fast_ready) throw Error(D4_ESTATE, "library tables are not on the device");
        if (max_h > (uint32_t)MAX_ROWS) throw Error(D4_EINVAL, "board taller than 64 rows");
        noalloc = true;
        heights.assign(1, 0);
        start.assign(1, 0);
        active.assign(1, 0);
        host_result.assign(1, 0);
        boards.assign(MAX_ROWS + 8, 0);
        occ_bits = 8 * (lib->maxh > 1 ? lib->maxh - 1 : 0);
        occ_mask = occ_bits ? (occ_bits >= 64 ? ~0ull : (1ull << occ_bits) - 1) : 0;
        used_mask = lib->used_bits >= 32 ? ~0u : (1u << lib->used_bits) - 1;
        if (occ_bits + lib->used_bits > 63) throw Error(D4_EINVAL, "state does not fit a 63-bit key");
        inst_shift = 64;
        st = na.st;
        ev0 = na.ev[0];
        ev1 = na.ev[1];
        evx0 = na.ev[2];
        evx1 = na.ev[3];
        evc0 = na.ev[4];
        bw_stride = max_h + 2;
        uint8_t *q = na.ptr;
        auto carve = [&](auto &buf, uint64_t n) {
            using T = std::remove_reference_t<decltype(*buf.p)>;
            const uint64_t bytes = (n * sizeof(T) + 255) & ~255ull;
            if ((uint64_t)(q - na.ptr) + bytes > kScratchBytes)
                throw Error(D4_EDEVICE, "internal error: scratch layout exceeds the array");
            buf.attach(reinterpret_cast<T *>(q), n);
            q += bytes;
        };
        carve(c32, 64);
        carve(c64, 64);
        carve(bw, (uint64_t)MAX_ROWS + 2);
        carve(hgt, 1);
        carve(dstart, 1);
        carve(result, 1);
        carve(result_save, 1);
        carve(blk, 2ull * COMPACT_MAX_BLOCKS);
        carve(spill, 2 * kNaSpill);
        carve(fk, kNaFront);
        carve(fc, kNaFront);
        carve(nk, kNaFront);
        carve(nc, kNaFront);
        carve(tk, 2 * kNaTable);
        carve(lv_a, (uint64_t)LVL_FIELDS * kNaLevel);
        carve(lv_b, (uint64_t)LVL_FIELDS * kNaLevel);
        carve(ch_k, kNaChild);
        carve(ch_c, kNaChild);
        // the scratch keeps the last call's contents: clear what starts at 0
        // (fill kernels, not memset nodes, in case a caller records them)
        launch_zero_u32(reinterpret_cast<uint32_t *>(bw.p), 2ull * bw_stride, st);
        launch_zero_u32(hgt.p, 1, st);
        launch_zero_u32(dstart.p, 1, st);
        launch_zero_u32(reinterpret_cast<uint32_t *>(result.p), 2, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }

    ~Engine() {
        (void)hipStreamSynchronize(st);
        if (noalloc) return;  // the stream and events belong to the scratch
        for (hipEvent_t e : {ev0, ev1, evx0, evx1, evc0})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : lv_ev) (void)hipEventDestroy(e);
        if (pin32) (void)hipHostFree(pin32);
        if (pin64) (void)hipHostFree(pin64);
        drop_graphs();
        if (ctl_h) (void)hipHostFree(ctl_h);
        if (st) (void)hipStreamDestroy(st);
    }

    // Queue an instance in a free slot; it joins the frontier at the next
    // step.
    void admit(uint32_t slot, const uint8_t *rows, uint32_t h) {
        if (slot >= S || active[slot]) throw Error(D4_EINVAL, "slot not free");
        if (h > max_h) throw Error(D4_EINVAL, "board taller than the engine's maximum");
        uint8_t *bd = &boards[(size_t)slot * (MAX_ROWS + 8)];
        std::memset(bd, 0, MAX_ROWS + 8);
        for (uint32_t r = 0; r < h; r++) {
            if (lib->W < 8 && (rows[r] >> lib->W))
                throw Error(D4_EINVAL, "board wider than the library's width");
            bd[r] = rows[r];
        }
        heights[slot] = h;
        active[slot] = 1;
        host_result[slot] = 0;
        pending.push_back(slot);
        n_active++;
    }

    bool idle() const { return n_active == 0; }

    // Device bytes held by this engine's buffers (high-water marks).
    uint64_t device_bytes() const {
        uint64_t b = 8 * (bw.cap + fk.cap + fc.cap + nk.cap + nc.cap + tk.cap + result.cap +
                          result_save.cap + c64.cap + e_desc.cap + e_slot.cap) +
                     4 * (hgt.cap + dstart.cap + c32.cap + e_parent.cap + e_child.cap +
                          slot_uid.cap + blk.cap);
        for (const DevLog &lg : logs)
            b += 8 * (lg.keys.cap + lg.desc.cap) + 4 * (lg.par.cap + lg.chi.cap);
        return b;
    }

    // rows still to be filled by a slot's instance (CudaSearcher::get_height)
    uint32_t height_of(uint32_t slot) const {
        if (!active[slot]) return 0;
        bool queued = std::find(pending.begin(), pending.end(), slot) != pending.end();
        return queued ? heights[slot] : heights[slot] - std::min(heights[slot], g - start[slot]);
    }

    ExtendArgs args(uint64_t tab_m, uint32_t e_cap, uint64_t begin, uint64_t count,
                    uint32_t split) {
        ExtendArgs a{};
        a.in_key = fk.p + begin;
        a.in_cnt = fc.p + begin;
        a.n_in = (uint32_t)count;
        a.rec_base = (uint32_t)begin;
        a.split = split;
        a.refill = lib->ext_refill;
        a.row = g;
        a.fit = lib->d_fit;
        a.placeinfo = lib->d_placeinfo;
        a.xavail = lib->d_xavail;
        a.fit_rows = lib->fit_rows;
        a.bw = bw.p;
        a.bw_stride = bw_stride;
        a.hgt = hgt.p;
        a.start = dstart.p;
        a.result = result.p;
        a.occ_bits = occ_bits;
        a.inst_shift = inst_shift;
        a.occ_mask = occ_mask;
        a.used_mask = used_mask;
        a.tab_key = tk.p;
        a.tab_cnt = tk.p + 1;  // interleaved {key, count} slots
        a.tab_size = tab_m;
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
        a.spill = mode == D4_MODE_COUNT && spill.p ? spill.p : nullptr;
        a.spill_cap = a.spill ? (uint32_t)std::min<uint64_t>(spill.cap / 2, kSpillCap) : 0;
        a.budget = lib->ext_budget;
        a.don_in = nullptr;
        a.n_don = 0;
        a.don_read = C32_DON;
        a.don_write = C32_DON;
        a.don_out = don_a.p;
        a.don_cap = (uint32_t)std::min<uint64_t>(don_a.cap / 2, lib->don_items);
        return a;
    }

    // Run the work items that lanes over their step budget handed over,
    // launch after launch, until none are left (after a few rounds the
    // budget is lifted so the chain ends).
    void run_donations(uint64_t tab_m, uint32_t e_cap, bool en, bool prune, bool count_only) {
        if (expand_kernel(en, count_only)) {
            run_expansions(tab_m, prune);
            return;
        }
        if (!lib->ext_budget) return;
        // Rounds are queued a few at a time without a host round trip: each
        // reads its item count from the counter pair the round before it
        // wrote, and the two pairs alternate.  A round lasts about as long
        // as its budget of dependent DFS steps on one lane, so follow-up
        // items get a short budget and split again quickly; after
        // kBudgetedRounds rounds the rest runs to completion.
        constexpr uint32_t kRoundsPerSync = 8;
        constexpr uint32_t kFollowBudget = 64;
        constexpr uint32_t kBudgetedRounds = 64;
        uint32_t rd = C32_DON;  // the pair that counted the items in don_a
        for (uint32_t round = 0;;) {
            uint32_t h[C32_N];
            HIPCHK(hipMemcpyAsync(h, c32.p, sizeof h, hipMemcpyDeviceToHost, st));
            sync_host();
            if (h[C32_BAD_ITEM]) throw Error(D4_EDEVICE, "internal error: bad donated work item");
            if (h[C32_OVERFLOW]) return;  // the step is rerun
            const uint32_t cap = (uint32_t)std::min<uint64_t>(don_a.cap / 2, lib->don_items);
            const uint32_t n =
                h[rd] <= cap ? h[rd] : (h[rd + 1] ? h[rd + 1] - 1 : cap);
            if (!n) return;
            for (uint32_t q = 0; q < kRoundsPerSync; q++, round++) {
                const uint32_t wr = rd == C32_DON ? C32_DON1 : C32_DON;
                launch_reset_donation_pair(c32.p, c64.p, wr, st);
                ExtendArgs a = args(tab_m, e_cap, 0, n_front, 0);
                a.don_in = don_a.p;
                a.n_don = q == 0 ? n : ~0u;
                a.don_read = rd;
                a.don_write = wr;
                a.don_out = don_b.p;
                a.don_cap = (uint32_t)std::min<uint64_t>(don_b.cap / 2, lib->don_items);
                a.budget = round < kBudgetedRounds ? kFollowBudget : 0;
                launch_extend(a, en, prune, count_only, lib->ext_threads, lib->n_cu, st);
                HIPCHK(hipGetLastError());
                don_a.swap(don_b);
                rd = wr;
            }
        }
    }

    // Count-mode row steps run k_expand unless the library selects the
    // depth-first kernel.
    bool expand_kernel(bool en, bool count_only) const {
        return !en && !count_only && lib->xp_kernel == D4_KERNEL_EXPAND;
    }
    bool level_kernel(bool en, bool count_only) const {
        return !en && !count_only && lib->xp_kernel == D4_KERNEL_LEVEL &&
               level_lds_bytes(lib->fit_rows) <= lib->lds_limit;
    }

    // ---- level kernels (count mode) -------------------------------------
    // Buffer budget: a quarter of the memory cap (16 GiB without a cap, at
    // most that).  A chunk of parents is sized from the per-parent partial
    // rows (widest level) and children of the last row step that had a
    // large frontier (a tiny frontier's first rows fan out by thousands per
    // parent and would understate the chunk afterwards); an overflow reruns
    // only the chunk that overflowed, with four times the estimate.
    uint64_t level_budget() const {
        if (noalloc) return 8ull * (2ull * lv_a.cap + 2ull * ch_k.cap);
        const uint64_t cap = mem_cap ? mem_cap / 4 : 16ull << 30;
        return std::max<uint64_t>(std::min<uint64_t>(cap, 16ull << 30), 1ull << 20);
    }
    // per-parent estimates at the start of a step
    // (pp_bound: the step's children bound per parent where it was computed;
    // the widest level has held well under a quarter of the children)
    double lv_pp0() const { return std::max({lv_ratio, 2.0, 0.25 * pp_bound, 1.1 * seen_lv}); }
    double ch_pp0() const { return std::max({ch_ratio, 4.0, pp_bound, 1.1 * seen_ch}); }
    uint64_t level_min_cap() const {
        return std::min<uint64_t>(std::max<uint64_t>(level_budget() / 1024, 4096), 1u << 20);
    }
    // parents per chunk: the buffers of a chunk fit in the budget
    uint64_t level_chunk(uint64_t n, double lv_pp, double ch_pp) const {
        const double pp = 8.0 * 1.5 * (2.0 * LVL_FIELDS * lv_pp + 2.0 * ch_pp);
        const uint64_t c = (uint64_t)((double)level_budget() / pp);
        return std::min<uint64_t>(std::max<uint64_t>(c, 1), n);
    }
    // capacities are powers of two, so that searches of similar size find
    // their buffers in the allocator's cache instead of allocating anew
    uint64_t level_cap(uint64_t c, double pp) const {
        return next_pow2(std::max<uint64_t>((uint64_t)((double)c * pp * 1.5), level_min_cap()));
    }
    // Device bytes the level buffers take for a row step of n parents.
    uint64_t level_bytes(uint64_t n) const {
        if (!n) return 0;
        const uint64_t c = level_chunk(n, lv_pp0(), ch_pp0());
        return 8ull * (2ull * LVL_FIELDS * level_cap(c, lv_pp0()) + 2ull * level_cap(c, ch_pp0()));
    }

    // Level-synchronous extend + dedup of the frontier (count mode).  Per
    // chunk of parents: one k_level launch per level (each reads its item
    // count on the device), then k_dedup merges the chunk's children into
    // the table; all chunks are queued with no host round trip.  A chunk
    // whose buffers overflow is rerun (chunks before it are complete; the
    // dedup of each completed chunk counts itself in c32[C32_LVDONE]).
    // Terminal records (instances that finish at this step) are counted by
    // one pass of their own, so a rerun chunk adds nothing twice.
    // The level kernels' and the dedup kernel's arguments for this step.
    void level_args(uint64_t tab_m, LevelArgs &a, DedupArgs &d) {
        if (!spill.ext) spill.reserve(2 * kSpillCap);
        a = LevelArgs{};
        a.fit = lib->d_fit;
        a.placeinfo = lib->d_placeinfo;
        a.xavail = lib->d_xavail;
        a.fit_rows = lib->fit_rows;
        a.bw = bw.p;
        a.bw_stride = bw_stride;
        a.hgt = hgt.p;
        a.start = dstart.p;
        a.result = result.p;
        a.row = g;
        a.occ_bits = occ_bits;
        a.inst_shift = inst_shift;
        a.used_mask = used_mask;
        a.occ_mask = occ_mask;
        a.c32 = c32.p;
        a.c64 = c64.p;
        d = DedupArgs{};
        d.bw = bw.p;
        d.bw_stride = bw_stride;
        d.start = dstart.p;
        d.row = g;
        d.occ_bits = occ_bits;
        d.inst_shift = inst_shift;
        d.occ_mask = occ_mask;
        d.seed_mask = lib->maxh >= 2 ? ~((1ull << (8 * (lib->maxh - 1))) - 1ull) : ~0ull;
        d.gcd = lib->gcd;
        d.min_size = lib->min_size;
        d.closed_rows = lib->maxh >= 1 ? lib->maxh - 1 : 0;
        d.area_ok = area_mask(lib->gcd, lib->min_size);
        d.n_inst = S;
        d.hgt = hgt.p;
        d.tab = tk.p;
        d.tab_size = tab_m;
        // linear probing at the loads the tables run at (up to about 75%
        // when a step is sized from the growth trend) has clusters of a few
        // hundred slots among billions of keys: a long probe limit keeps
        // those keys out of the spill list, whose contents cost a whole
        // table growth
        d.max_probe = 4096;
        d.spill = spill.p;
        d.spill_cap = (uint32_t)std::min<uint64_t>(spill.cap / 2, kSpillCap);
        d.c32 = c32.p;
        d.c64 = c64.p;
        // (the level kernels prune with the dedup's fields when asked to)
        a.seed_mask = d.seed_mask;
        a.gcd = d.gcd;
        a.min_size = d.min_size;
        a.closed_rows = d.closed_rows;
        a.area_ok = d.area_ok;
        a.n_win0 = S <= 1024 && level_lds_bytes(lib->fit_rows, 0, S) <= lib->lds_limit ? S : 0;
        a.n_win =
            S <= 1024 && level_lds_bytes(lib->fit_rows, S, a.n_win0) <= lib->lds_limit ? S : 0;
        a.stage = level_lds_bytes(lib->fit_rows, a.n_win, a.n_win0, true) <= lib->lds_limit;
    }

    // Queues the terminal pass when an instance finishes at this step.
    void queue_terminal() { queue_terminal(fk.p, fc.p, n_front); }
    void queue_terminal(const uint64_t *k, const uint64_t *c, uint64_t n) {
        bool any_term = false;
        for (uint32_t i = 0; i < S; i++)
            if (active[i] && start[i] + heights[i] == g) any_term = true;
        if (any_term && n)
            launch_terminal(k, c, n, g, dstart.p, hgt.p, inst_shift, result.p, lib->n_cu, st);
    }

    // Sizes the level buffers for chunks of `chunk` parents.
    void level_buffers(uint64_t chunk, double lv_pp, double ch_pp, LevelArgs &a, DedupArgs &d) {
        lv_cap = level_cap(chunk, lv_pp);
        ch_cap = level_cap(chunk, ch_pp);
        if (noalloc) {
            lv_cap = std::min<uint64_t>(lv_cap, lv_a.cap / LVL_FIELDS);
            ch_cap = std::min<uint64_t>(ch_cap, ch_k.cap);
        }
        lv_a.reserve(LVL_FIELDS * lv_cap);
        lv_b.reserve(LVL_FIELDS * lv_cap);
        ch_k.reserve(ch_cap);
        ch_c.reserve(ch_cap);
        a.ch_key = ch_k.p;
        a.ch_cnt = ch_c.p;
        d.ch_key = ch_k.p;
        d.ch_cnt = ch_c.p;
        a.ch_cap = d.ch_cap = ch_cap;
        a.in_cap = a.out_cap = lv_cap;
    }

    // Queues chunks b0, b0 + chunk, ... of the frontier: per chunk the level
    // launches and the dedup, with events around them when timed.
    // (pa set: the chunk -- the whole frontier -- is deduped by partitions
    // into pa's next frontier instead of the table)
    uint32_t queue_levels(LevelArgs &a, const DedupArgs &d, uint64_t b0, uint64_t chunk,
                          bool prune, bool timed, const PartArgs *pa = nullptr) {
        return queue_levels(a, d, fk.p, fc.p, n_front, b0, chunk, prune, timed, pa);
    }
    uint32_t queue_levels(LevelArgs &a, const DedupArgs &d, const uint64_t *ik,
                          const uint64_t *ic, uint64_t n_in, uint64_t b0, uint64_t chunk,
                          bool prune, bool timed, const PartArgs *pa = nullptr) {
        const uint32_t nch = (uint32_t)((n_in - b0 + chunk - 1) / chunk);
        while (timed && lv_ev.size() < 2 * (size_t)nch + 1) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            lv_ev.push_back(e);
        }
        // the connectivity prune runs in the level kernels, before a child
        // is written, or in the dedup
        level_prune = level_prunes(a);
        a.prune = prune && level_prune;
        const bool dprune = prune && !a.prune;
        launch_reset_levels(c32.p, c64.p, true, st);
        if (timed) HIPCHK(hipEventRecord(lv_ev[0], st));
        for (uint32_t j = 0; j < nch; j++) {
            const uint64_t b = b0 + (uint64_t)j * chunk;
            // (with the partitioned dedup the chunks' children accumulate and
            // meet in one dedup after the last chunk)
            if (j) launch_reset_levels(c32.p, c64.p, false, st, nullptr, pa != nullptr);
            a.in_key = ik + b;
            a.in_cnt = ic + b;
            a.n_rec = (uint32_t)std::min<uint64_t>(chunk, n_in - b);
            for (uint32_t L = 0; L < LVL_MAX; L++) {
                a.level = L;
                a.in_st = (L & 1) ? lv_a.p : lv_b.p;
                a.out_st = (L & 1) ? lv_b.p : lv_a.p;
                launch_level(a, lib->n_cu, st);
            }
            if (timed) HIPCHK(hipEventRecord(lv_ev[2 * j + 1], st));
            if (!pa) launch_dedup(d, dprune, lib->n_cu, st);
            else if (j + 1 == nch) launch_part_dedup(*pa, dprune, lib->n_cu, st);
            if (timed) HIPCHK(hipEventRecord(lv_ev[2 * j + 2], st));
        }
        return nch;
    }

    // Kernel times of the chunks queued last (after the stream is synced).
    void level_times(uint32_t nch) {
        for (uint32_t j = 0; j < nch; j++) {
            float lms = 0, dms = 0;
            HIPCHK(hipEventElapsedTime(&lms, j ? lv_ev[2 * j] : lv_ev[0], lv_ev[2 * j + 1]));
            HIPCHK(hipEventElapsedTime(&dms, lv_ev[2 * j + 1], lv_ev[2 * j + 2]));
            stats.extend_ms += lms;
            stats.dedup_ms += dms;
            stats.dedup_launches++;
            stats.extend_launches++;
        }
    }

    // Per-parent sizes of a completed chunk of `c` parents (its counters are
    // still set) for the next steps' chunking.
    // (c_ch: the parents whose children C64_CHW counts, when the chunks'
    // children accumulated; 0: c)
    void level_sizes(const unsigned long long *h64, uint64_t c, uint64_t c_ch = 0) {
        if (!c) return;
        uint64_t m = 0;
        for (uint32_t L = 0; L < LVL_MAX; L++) m = std::max<uint64_t>(m, h64[C64_LVL + L]);
        const double lv = (double)m / (double)c,
                     ch = (double)h64[C64_CHW] / (double)(c_ch ? c_ch : c);
        lib->note_sizes(stats.rows_done, lv, ch);
        if (c < 4096) return;
        lv_ratio = lv;
        ch_ratio = ch;
    }

    void run_levels(uint64_t tab_m, bool prune) {
        run_levels(fk.p, fc.p, n_front, tab_m, prune, 0, 0, true);
    }
    // The level launches and dedups over n_in parents at ik / ic, in chunks
    // that fit the level buffers, rerunning a chunk whose buffers overflowed.
    // Only children whose key hash has (hash & part_mask) == part_id enter
    // the table (a streamed row step's partition); terminal: queue the
    // terminal pass over these parents first.
    void run_levels(const uint64_t *ik, const uint64_t *ic, uint64_t n_in, uint64_t tab_m,
                    bool prune, uint32_t part_mask, uint32_t part_id, bool terminal) {
        LevelArgs a;
        DedupArgs d;
        level_args(tab_m, a, d);
        d.part_mask = part_mask;
        d.part_id = part_id;
        if (terminal) queue_terminal(ik, ic, n_in);
        uint64_t b0 = 0;
        uint32_t reruns = 0;
        double lv_pp = lv_pp0(), ch_pp = ch_pp0();
        for (;;) {
            const uint64_t chunk = level_chunk(n_in - b0, lv_pp, ch_pp);
            level_buffers(chunk, lv_pp, ch_pp, a, d);
            const bool timed = !noalloc;  // (no events are made for the scratch's chunks)
            const uint32_t nch = queue_levels(a, d, ik, ic, n_in, b0, chunk, prune, timed);
            HIPCHK(hipGetLastError());
            uint32_t h32[C32_N];
            unsigned long long h64[C64_N];
            HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
            sync_host();
            const uint32_t done = std::min<uint32_t>(h32[C32_LVDONE], nch);
            if (timed) level_times(nch);
            if (h32[C32_OVERFLOW]) return;  // the table: the row step reruns larger
            if (done == nch && !h32[C32_XOVF])
                level_sizes(h64, std::min<uint64_t>(chunk, n_in - (b0 + (uint64_t)(nch - 1) * chunk)));
            if (trace)
                fprintf(stderr,
                        "levels step %u: parents %llu from %llu, chunk %llu, caps %llu / %llu, "
                        "done %u of %u, overflow %u, children %llu, level rows %llu %llu %llu %llu "
                        "%llu %llu %llu %llu\n",
                        g, (unsigned long long)n_in, (unsigned long long)b0,
                        (unsigned long long)chunk, (unsigned long long)lv_cap,
                        (unsigned long long)ch_cap, done, nch, h32[C32_XOVF], h64[C64_CHW],
                        h64[C64_LVL], h64[C64_LVL + 1], h64[C64_LVL + 2], h64[C64_LVL + 3],
                        h64[C64_LVL + 4], h64[C64_LVL + 5], h64[C64_LVL + 6], h64[C64_LVL + 7]);
            if (!h32[C32_XOVF]) break;
            // rerun from the chunk that overflowed, with four times the room
            // per parent for the stream that ran out (a counter that passed
            // its capacity)
            stats.chunk_reruns++;
            b0 += (uint64_t)done * chunk;
            const uint64_t c_fail = std::min<uint64_t>(chunk, n_in - b0);
            bool lv_over = false;
            for (uint32_t L = 0; L < LVL_MAX; L++) lv_over |= h64[C64_LVL + L] > lv_cap;
            if (lv_over) lv_pp = 4.0 * std::max(lv_pp, (double)lv_cap / (double)c_fail);
            if (h64[C64_CHW] > ch_cap || !lv_over)
                ch_pp = 4.0 * std::max(ch_pp, (double)ch_cap / (double)c_fail);
            if (++reruns > 32)
                throw Error(D4_ERANGE, noalloc ? "allocation-free search: a row step's partial "
                                                 "rows exceed the scratch"
                                               : "a row step's partial rows outgrow every buffer");
        }
    }

    // k_expand's follow-up launches: the partial-row states spilled by the
    // launch before (a full stack, or a wave's tail), eight launches per host
    // round trip, each reading its item count on the device, until a round
    // spills nothing.  Tail spills stop after kBudgetedRounds rounds, so the
    // chain ends (a full stack still spills).
    void run_expansions(uint64_t tab_m, bool prune) {
        constexpr uint32_t kRoundsPerSync = 8;
        constexpr uint32_t kBudgetedRounds = 64;
        uint32_t rd = C32_DON;
        for (uint32_t round = 0;;) {
            uint32_t h[C32_N];
            HIPCHK(hipMemcpyAsync(h, c32.p, sizeof h, hipMemcpyDeviceToHost, st));
            sync_host();
            if (h[C32_OVERFLOW] || h[C32_XOVF]) return;  // the step is rerun
            const uint32_t cap = (uint32_t)std::min<uint64_t>(don_a.cap / EXPAND_WORDS, xp_cap);
            const uint32_t n = h[rd] <= cap ? h[rd] : (h[rd + 1] ? h[rd + 1] - 1 : cap);
            if (!n) return;
            for (uint32_t q = 0; q < kRoundsPerSync; q++, round++) {
                const uint32_t wr = rd == C32_DON ? C32_DON1 : C32_DON;
                launch_reset_donation_pair(c32.p, c64.p, wr, st);
                ExtendArgs a = args(tab_m, 0, 0, n_front, 0);
                a.don_in = don_a.p;
                a.n_don = q == 0 ? n : ~0u;
                a.don_read = rd;
                a.don_write = wr;
                a.don_out = don_b.p;
                a.don_cap = (uint32_t)std::min<uint64_t>(don_b.cap / EXPAND_WORDS, xp_cap);
                a.budget = round < kBudgetedRounds ? lib->xp_budget : 0;
                launch_expand(a, prune, lib->ext_threads, lib->n_cu, st);
                HIPCHK(hipGetLastError());
                don_a.swap(don_b);
                rd = wr;
            }
        }
    }

    // Move queued instances into the frontier: their board windows, height,
    // start step and a zero count go to the device, and a root state (empty
    // occupancy, count 1) is appended to the frontier.
    void upload_pending() {
        if (pending.empty()) return;
        std::vector<uint64_t> k0, c0;
        for (uint32_t slot : pending) {
            start[slot] = g;
            k0.push_back(inst_shift < 64 ? (uint64_t)slot << inst_shift : 0);
            c0.push_back(1);
        }
        // board windows and zeroed counts, one copy per run of consecutive
        // slots (a batch admits all its boards as one run)
        std::vector<uint32_t> order(pending.begin(), pending.end());
        std::sort(order.begin(), order.end());
        // (all copies are queued, then one wait: the host buffers live until it)
        std::deque<std::vector<uint64_t>> wins;
        std::deque<std::vector<unsigned long long>> zeros;
        auto copy_async = [&](void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
            HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
        };
        for (size_t i = 0; i < order.size();) {
            size_t j = i + 1;
            while (j < order.size() && order[j] == order[j - 1] + 1) j++;
            const uint32_t first = order[i], run = (uint32_t)(j - i);
            std::vector<uint64_t> &win = wins.emplace_back((size_t)run * bw_stride, 0);
            for (uint32_t q = 0; q < run; q++) {
                const uint8_t *bd = &boards[(size_t)(first + q) * (MAX_ROWS + 8)];
                for (uint32_t r = 0; r < bw_stride; r++) {
                    uint64_t w = 0;
                    for (uint32_t b = 0; b < 8 && r + b < (uint32_t)MAX_ROWS + 8; b++)
                        w |= (uint64_t)bd[r + b] << (8 * b);
                    win[(size_t)q * bw_stride + r] = w;
                }
            }
            copy_async(bw.p + (size_t)first * bw_stride, win.data(), win.size() * 8,
                       hipMemcpyHostToDevice);
            std::vector<unsigned long long> &zero = zeros.emplace_back(run, 0);
            copy_async(result.p + first, zero.data(), (size_t)run * 8, hipMemcpyHostToDevice);
            i = j;
        }
        copy_async(hgt.p, heights.data(), (size_t)S * 4, hipMemcpyHostToDevice);
        copy_async(dstart.p, start.data(), (size_t)S * 4, hipMemcpyHostToDevice);
        const uint64_t n = n_front + k0.size();
        if (n > fk.cap || n > fc.cap) {  // grow, keeping the current frontier
            DevBuf<uint64_t> k2, c2;
            k2.reserve(n);
            c2.reserve(n);
            if (n_front) {
                copy_async(k2.p, fk.p, n_front * 8, hipMemcpyDeviceToDevice);
                copy_async(c2.p, fc.p, n_front * 8, hipMemcpyDeviceToDevice);
            }
            // (the old buffers go back to the allocator's cache only after
            // the copies out of them have ended)
            sync_host();
            fk.swap(k2);
            fc.swap(c2);
        }
        copy_async(fk.p + n_front, k0.data(), k0.size() * 8, hipMemcpyHostToDevice);
        copy_async(fc.p + n_front, c0.data(), c0.size() * 8, hipMemcpyHostToDevice);
        sync_host();
        n_front = n;
        new_roots = k0.size();
        stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, n_front);
        pending.clear();
    }

    // One global step: admitted instances join, every instance advances one
    // row (search_GPU), and those that reach their height finish.
    void step() {
        completed.clear();
        upload_pending();
        if (!n_active) throw Error(D4_ESTATE, "no instance left to search");
        row_step();
        g++;
        take_sample();
        g--;
        bool any = false;
        for (uint32_t i = 0; i < S; i++)
            if (active[i] && start[i] + heights[i] == g) any = true;
        if (any) {
            copy(host_result.data(), result.p, (size_t)S * 8, hipMemcpyDeviceToHost);
            for (uint32_t i = 0; i < S; i++)
                if (active[i] && start[i] + heights[i] == g) {
                    completed.push_back(i);
                    active[i] = 0;
                    n_active--;
                }
        }
        g++;
    }

    void run() {
        if (mem_cap && mode == D4_MODE_COUNT) run_bounded();
        else
            while (n_active) step();
    }

    void drop_graphs() {
        for (hipGraphExec_t &x : gx)
            if (x) {
                (void)hipGraphExecDestroy(x);
                x = nullptr;
            }
        g_sig.clear();
    }

    // What a recorded step depends on: every buffer address and capacity it
    // was recorded with (a buffer the general path grows moves, and the
    // graphs are recorded again).
    std::vector<uint64_t> graph_sig(bool prune) const {
        std::vector<uint64_t> v;
        for (const void *q : {(const void *)tk.p, (const void *)lv_a.p, (const void *)lv_b.p,
                              (const void *)ch_k.p, (const void *)ch_c.p, (const void *)spill.p,
                              (const void *)result.p, (const void *)result_save.p,
                              (const void *)bw.p, (const void *)hgt.p, (const void *)dstart.p,
                              (const void *)c32.p, (const void *)c64.p, (const void *)blk.p,
                              (const void *)gk[0].p, (const void *)gk[1].p,
                              (const void *)gc[0].p, (const void *)gc[1].p,
                              (const void *)ctl_d.p, (const void *)pbuf.p, (const void *)sbuf.p,
                              (const void *)pmask.p, (const void *)phist.p,
                              (const void *)phist2.p, (const void *)pbsum.p})
            v.push_back((uint64_t)(uintptr_t)q);
        for (uint64_t x : {tk.cap, lv_a.cap, lv_b.cap, ch_k.cap, ch_c.cap, spill.cap, g_cap, g_lv,
                           g_ch, (uint64_t)S, (uint64_t)prune, (uint64_t)bw_stride,
                           (uint64_t)g_part})
            v.push_back(x);
        return v;
    }

    // Records the row step that reads frontier buffer p and writes 1 - p:
    // result save, table fill, counter resets, terminal pass, the eight level
    // launches, the dedup, compaction and the close kernel, all sized from
    // the control block (launch arguments are the capacities).
    void record_step(uint32_t p, bool prune) {
        LevelArgs a;
        DedupArgs d;
        level_args(g_cap, a, d);
        a.ch_key = ch_k.p;
        a.ch_cnt = ch_c.p;
        d.ch_key = ch_k.p;
        d.ch_cnt = ch_c.p;
        a.ch_cap = d.ch_cap = g_ch;
        a.in_cap = a.out_cap = g_lv;
        a.ctl = d.ctl = ctl_d.p;
        a.in_key = gk[p].p;
        a.in_cnt = gc[p].p;
        a.n_rec = (uint32_t)std::min<uint64_t>(g_cap, 0xffffffffull);
        level_prune = level_prunes(a);
        a.prune = prune && level_prune;
        const bool dprune = prune && !a.prune;
        const RowCtl *ctl = ctl_d.p;
        // (the partitioned dedup writes the next frontier buffer itself)
        const bool part = g_part;
        PartArgs pa{};
        if (part) part_args(d, g_ch, gk[1 - p].p, gc[1 - p].p, g_cap, pa);
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launch_copy_u64(result_save.p, result.p, S, st, ctl);
        if (!part) launch_fill_table(tk.p, tk.p + 1, g_cap, st, ctl);
        launch_reset_counters(c32.p, c64.p, st, ctl);
        launch_terminal(gk[p].p, gc[p].p, g_cap, g, dstart.p, hgt.p, inst_shift, result.p,
                        lib->n_cu, st, ctl);
        launch_reset_levels(c32.p, c64.p, true, st, ctl);
        for (uint32_t L = 0; L < LVL_MAX; L++) {
            a.level = L;
            a.in_st = (L & 1) ? lv_a.p : lv_b.p;
            a.out_st = (L & 1) ? lv_b.p : lv_a.p;
            launch_level(a, lib->n_cu, st);
        }
        if (part) {
            launch_part_dedup(pa, dprune, lib->n_cu, st);
        } else {
            launch_dedup(d, dprune, lib->n_cu, st);
            launch_compact(tk.p, tk.p + 1, g_cap, gk[1 - p].p, gc[1 - p].p, g_cap, c32.p, blk.p,
                           nullptr, st, ctl);
        }
        launch_row_close(ctl_d.p, c32.p, c64.p, st);
        const hipError_t launch_err = hipGetLastError();
        hipGraph_t graph = nullptr;
        const hipError_t cap_err = hipStreamEndCapture(st, &graph);
        HIPCHK(launch_err);
        HIPCHK(cap_err);
        const hipError_t inst_err = hipGraphInstantiate(&gx[p], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(inst_err);
        // (uploaded now, so that the first replay does not pay for it)
        HIPCHK(hipGraphUpload(gx[p], st));
    }

    // Up to kmax count-mode row steps replayed from the recorded graphs, with
    // one host round trip for all of them.  tab0: the first step's table
    // slots (0: from the unique states per parent of the step before).
    // Returns the steps completed, their bookkeeping done (the caller moves
    // g); a step that failed leaves the frontier before it in fk / fc, its
    // terminal counts undone, and its flags in g_fail.  Returns 0 with g_fail
    // clear where replay does not apply.
    uint32_t graph_rows(uint32_t kmax, uint64_t tab0, bool prune) {
        g_fail = 0;
        if (lib->step_mode != D4_STEP_GRAPH || mode != D4_MODE_COUNT || noalloc || !n_front ||
            !kmax || !level_kernel(false, false) || sample_h || n_front > 0xffffffffull)
            return 0;
        if (!tab0)
            tab0 = round_slots((uint64_t)(1.4 * table_scale *
                                          ((double)n_front * std::max(grow_unique_raw, 0.01) *
                                               1.1 + 4096.0)));
        // Capacities for every step of the run: the table and frontier
        // buffers with room for eight times the first step's table (capped;
        // the frontier buffers hold the current frontier too), and level
        // buffers for that many parents at the widest sizes per parent seen
        // so far (this search's and the library's for the rows ahead).  A
        // step that outgrows them fails and reruns; spare room costs memory
        // only, as the kernels read the step's sizes from the control block.
        if (2 * tab0 > kGraphMaxSlots || n_front > kGraphMaxSlots) return 0;
        const uint64_t cap = std::max<uint64_t>(
            g_cap, std::min<uint64_t>(kGraphMaxSlots,
                                      next_pow2(std::max<uint64_t>({8 * tab0, n_front, 1u << 16}))));
        // (the first step's own sizes per parent include its children bound,
        // which applies to its parents only)
        double lpp = std::max(lv_ratio, 2.0), cpp = std::max(ch_ratio, 4.0);
        for (uint64_t r = stats.rows_done; r < stats.rows_done + kmax && r < lib->kSizeRows; r++) {
            double a = 0, b = 0;
            lib->seen_sizes(r, a, b);
            lpp = std::max(lpp, 1.1 * a);
            cpp = std::max(cpp, 1.1 * b);
        }
        const uint64_t par = std::min<uint64_t>(cap, std::max<uint64_t>(8 * n_front, 1u << 16));
        const uint64_t lv = std::max<uint64_t>(
            {g_lv, level_cap(par, lpp), level_cap(n_front, lv_pp0())});
        const uint64_t ch = std::max<uint64_t>(
            {g_ch, level_cap(par, cpp), level_cap(n_front, ch_pp0())});
        if (2ull * 8 * LVL_FIELDS * lv + 16ull * ch > level_budget()) return 0;
        if (mem_cap && 16ull * cap + 32ull * cap + 2ull * 8 * LVL_FIELDS * lv + 16ull * ch +
                               16ull * n_front + parked_bytes > mem_cap)
            return 0;
        g_cap = cap;
        g_lv = lv;
        g_ch = ch;
        g_part = part_dedup_fits(g_ch, n_front, g_cap, (double)g_ch);
        if (!g_part) tk.reserve(2 * g_cap);
        for (int q = 0; q < 2; q++) {
            gk[q].reserve(g_cap);
            gc[q].reserve(g_cap);
        }
        lv_a.reserve(LVL_FIELDS * g_lv);
        lv_b.reserve(LVL_FIELDS * g_lv);
        ch_k.reserve(g_ch);
        ch_c.reserve(g_ch);
        lv_cap = g_lv;
        ch_cap = g_ch;
        if (g_part) {
            const uint64_t nm = std::min<uint64_t>(g_ch, PART_MAX_CHILDREN);
            uint64_t h1 = 0, h2 = 0;
            part_hist_entries(nm, part_tile_min(), part_lp_max(), h1, h2);
            pbuf.reserve(2 * nm);
            sbuf.reserve(2 * nm);
            pmask.reserve(part_mask_words(nm));
            phist.reserve(h1);
            phist2.reserve(h2);
            pbsum.reserve(part_bsum_entries());
        } else {
            blk.reserve(2 * COMPACT_MAX_BLOCKS);
        }
        if (!spill.ext) spill.reserve(2 * kSpillCap);
        result_save.reserve(S);
        ctl_d.reserve(1);
        if (!ctl_h) HIPCHK(hipHostMalloc((void **)&ctl_h, sizeof(RowCtl)));
        const std::vector<uint64_t> sig = graph_sig(prune);
        if (sig != g_sig) {
            drop_graphs();
            record_step(0, prune);
            record_step(1, prune);
            g_sig = sig;
        }
        const uint32_t k = std::min<uint32_t>(kmax, CTL_LOG);
        HIPCHK(hipMemcpyAsync(gk[0].p, fk.p, n_front * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(gc[0].p, fc.p, n_front * 8, hipMemcpyDeviceToDevice, st));
        ctl_h->n_front = n_front;
        ctl_h->tab_m = std::min(tab0, g_cap);
        ctl_h->tab_cap = g_cap;
        ctl_h->row = g;
        ctl_h->ratio = stats.rows_done ? grow_unique_raw : 0.0;
        ctl_h->fail = 0;
        ctl_h->done = 0;
        HIPCHK(hipMemcpyAsync(ctl_d.p, ctl_h, offsetof(RowCtl, log), hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(ev0, st));
        for (uint32_t j = 0; j < k; j++) HIPCHK(hipGraphLaunch(gx[j & 1], st));
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipMemcpyAsync(ctl_h, ctl_d.p, sizeof(RowCtl), hipMemcpyDeviceToHost, st));
        sync_host();
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.gpu_ms += ms;
        stats.graph_replays += k;
        const uint32_t done = std::min(ctl_h->done, k);
        g_fail = ctl_h->fail;
        for (uint32_t j = 0; j < done; j++) {
            const StepLog &l = ctl_h->log[j];
            const uint64_t n = l.n_front, u = l.fresh;
            check_row_states(u);
            check_compacted(l.out, u);
            stats.children += l.children;
            stats.pruned += l.pruned;
            stats.records_extended += n;
            stats.unique += u;
            stats.extend_bytes += 16ull * n + 16ull * l.children;
            stats.dedup_bytes += 16ull * (l.children + (level_prune ? 0 : l.pruned)) + 16ull * u;
            if (g_part) {
                stats.part_steps++;
            } else {
                stats.compact_bytes += 8ull * l.tab_m + 24ull * u;
                stats.compact_launches++;
            }
            stats.extend_launches++;
            stats.dedup_launches++;
            if (n) {
                prev_unique_raw = stats.rows_done ? grow_unique_raw : 0.0;
                grow_unique = std::max(1.0, (double)u / (double)n);
                grow_unique_raw = std::max(0.01, (double)u / (double)n);
                grow_children = std::max(1.0, (double)l.children / (double)n);
                lib->note_sizes(stats.rows_done, (double)l.lv_max / (double)n,
                                (double)l.chw / (double)n);
                if (n >= 4096) {
                    lv_ratio = (double)l.lv_max / (double)n;
                    ch_ratio = (double)l.chw / (double)n;
                }
            }
            stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, u);
            stats.rows_done++;
            if (trace)
                fprintf(stderr, "graph step %u: parents %llu, table %llu, children %llu, new %llu\n",
                        g + j, (unsigned long long)n, (unsigned long long)l.tab_m,
                        (unsigned long long)l.chw, (unsigned long long)u);
        }
        // the frontier after the last completed step (the input of a failed one)
        const uint32_t q = done & 1;
        n_front = ctl_h->n_front;
        fk.reserve(std::max<uint64_t>(n_front, 1));
        fc.reserve(std::max<uint64_t>(n_front, 1));
        HIPCHK(hipMemcpyAsync(fk.p, gk[q].p, n_front * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(fc.p, gc[q].p, n_front * 8, hipMemcpyDeviceToDevice, st));
        if (g_fail) {
            // undo the failed step's terminal counts (its save is the last one)
            stats.fused_fallbacks++;
            HIPCHK(hipMemcpyAsync(result.p, result_save.p, (size_t)S * 8,
                                  hipMemcpyDeviceToDevice, st));
            if (trace)
                fprintf(stderr, "graph step %u: failed (flags %u), rerun on the general path\n",
                        g + done, g_fail);
        }
        if (done) new_roots = 0;
        return done;
    }

    // Where count-mode row steps run the connectivity prune.  In the level
    // kernels (before a child is written, so that the dedup reads survivors
    // only) when each wave's staging area fits in LDS next to the tables:
    // the wave's completed children are gathered there and tested 64 at a
    // time.  Otherwise in the dedup: tested on the completing lanes alone
    // (the rest of the wave idle), the flood fill costs the level kernels
    // more than the dedup saves.  D4_LEVEL_PRUNE=0 keeps it in the dedup,
    // =1 puts it in the level kernels either way.
    int level_prune_mode = [] {
        const char *e = std::getenv("D4_LEVEL_PRUNE");
        return e && *e == '0' ? 0 : e && *e == '1' ? 1 : -1;
    }();
    bool level_prune = false;  // (the last row step's choice)
    bool level_prunes(const LevelArgs &a) const {
        return level_prune_mode == 1 || (level_prune_mode < 0 && a.stage);
    }

    // Whether the step's children (ch_cap of them) go through the
    // partitioned dedup: the library asks for it, the children fit its
    // buckets, and its buffers (children in bucket order, a next frontier
    // with room for every child, the histogram) fit the memory cap.
    // (rows expected to have fewer than part_min children keep the table: the
    // bucket chain's dozen launches cost more than it saves there; 2^20 by
    // default, D4_PART_MIN_CHILDREN overrides it, so tests can send small
    // rows through the buckets)
    double part_min = [] {
        const char *v = getenv("D4_PART_MIN_CHILDREN");
        return v ? std::max(0.0, atof(v)) : (double)(1u << 20);
    }();
    // Limits of the partitioned dedup below its built-in ones, which reach
    // its recovery paths on small searches (tests): the children its
    // buffers take (D4_PART_MAX_CHILDREN; a row with more falls back to the
    // table) and the next-frontier states the merge is given room for
    // (D4_PART_OUT_CAP; finding more, it merges again into a frontier of
    // the size it counted).
    uint64_t part_n_cap = [] {
        const char *v = getenv("D4_PART_MAX_CHILDREN");
        const unsigned long long x = v ? strtoull(v, nullptr, 10) : 0ull;
        return x ? std::min<uint64_t>(x, PART_MAX_CHILDREN) : PART_MAX_CHILDREN;
    }();
    uint64_t part_out_cap = [] {
        const char *v = getenv("D4_PART_OUT_CAP");
        return (uint64_t)(v ? strtoull(v, nullptr, 10) : 0ull);
    }();
    bool part_dedup_fits(uint64_t chc, uint64_t nf, uint64_t out_cap, double expect) const {
        if (lib->dedup_mode != D4_DEDUP_PART || noalloc) return false;
        if (expect < part_min) return false;
        if (!mem_cap) return true;
        const uint64_t nm = std::min<uint64_t>(chc, PART_MAX_CHILDREN);
        uint64_t h1 = 0, h2 = 0;
        part_hist_entries(nm, part_tile_min(), part_lp_max(), h1, h2);
        return 32ull * nm + nm / 8 + 4ull * (h1 + h2) + 16ull * out_cap + 16ull * nf +
                   level_bytes(nf) + parked_bytes <= mem_cap;
    }
    // Whether a row step too large for one chunk of level buffers can still
    // be fused: its chunks' children accumulate in one buffer for the row and
    // meet in the partitioned dedup, which must apply and whose buffers
    // (children, survivors twice, histograms, the next frontier) must fit the
    // cap, or half the device's free memory without one.
    bool fused_chunks_ok(uint64_t tab_m) const {
        const double ch_pp = ch_pp0();
        const uint64_t chr = level_cap(n_front, ch_pp);
        if (chr > PART_MAX_CHILDREN) return false;
        if (!part_dedup_fits(chr, n_front, tab_m, ch_pp * (double)n_front)) return false;
        const uint64_t c = level_chunk(n_front, lv_pp0(), ch_pp);
        const uint64_t bytes = 48ull * chr + 16ull * tab_m + 16ull * n_front + parked_bytes +
                               8ull * 2 * LVL_FIELDS * level_cap(c, lv_pp0());
        const uint64_t cap = mem_cap ? mem_cap : available_device_bytes() / 2;
        return bytes <= cap;
    }

    // The partitioned dedup's arguments (pruning as the table dedup's d).
    void part_args(const DedupArgs &d, uint64_t chc, uint64_t *ok, uint64_t *oc, uint64_t out_cap,
                   PartArgs &pa) {
        const uint64_t nm = std::min<uint64_t>(chc, PART_MAX_CHILDREN);
        // (the bucket shape is read once: the buffers are sized for it)
        const uint64_t tile_min = part_tile_min();
        const uint32_t lp_max = part_lp_max();
        uint64_t h1 = 0, h2 = 0;
        part_hist_entries(nm, tile_min, lp_max, h1, h2);
        pbuf.reserve(2 * nm);
        sbuf.reserve(2 * nm);
        pmask.reserve(part_mask_words(nm));
        phist.reserve(h1);
        phist2.reserve(h2);
        pbsum.reserve(part_bsum_entries());
        pa = PartArgs{};
        // (the second radix pass writes the survivors back into the
        // children buffers, which the level kernels wrote)
        pa.ch_key = ch_k.p;
        pa.ch_cnt = ch_c.p;
        pa.ch_cap = chc;
        pa.n_max = std::min<uint64_t>(nm, part_n_cap);
        pa.tile_min = tile_min;
        pa.lp_max = lp_max;
        pa.slots = part_slots();
        pa.mask = pmask.p;
        pa.pbuf = pbuf.p;
        pa.sbuf = sbuf.p;
        pa.hist = phist.p;
        pa.hist2 = phist2.p;
        pa.bsum = pbsum.p;
        pa.out_key = ok;
        pa.out_cnt = oc;
        pa.out_cap = part_out_cap ? std::min<uint64_t>(out_cap, part_out_cap) : out_cap;
        pa.bw = d.bw;
        pa.bw_stride = d.bw_stride;
        pa.start = d.start;
        pa.row = d.row;
        pa.inst_shift = d.inst_shift;
        pa.occ_mask = d.occ_mask;
        pa.seed_mask = d.seed_mask;
        pa.gcd = d.gcd;
        pa.min_size = d.min_size;
        pa.closed_rows = d.closed_rows;
        pa.area_ok = d.area_ok;
        pa.n_inst = d.n_inst;
        pa.c32 = d.c32;
        pa.c64 = d.c64;
        pa.ctl = d.ctl;
    }

    // The count-mode row step as one queued chain with one host round trip
    // (level kernels, the frontier in a single chunk): table fill, counter
    // resets, terminal pass, the level launches, the dedup, and compaction
    // into next-frontier buffers sized to the table (the unique count is not
    // known before the chain ends), then the counters.  Returns false, with
    // the terminal counts restored, when a buffer overflowed or children
    // spilled: the general path then runs the step.
    bool fused_step(uint64_t tab_m, bool prune, uint32_t *h32, unsigned long long *h64) {
        if (!pin32) {
            HIPCHK(hipHostMalloc((void **)&pin32, sizeof(uint32_t) * C32_N));
            HIPCHK(hipHostMalloc((void **)&pin64, sizeof(unsigned long long) * C64_N));
        }
        LevelArgs a;
        DedupArgs d;
        level_args(tab_m, a, d);
        const double lv_pp = lv_pp0(), ch_pp = ch_pp0();
        // a frontier too large for one chunk of level buffers runs in chunks
        // whose children accumulate in one buffer for the row (partitioned
        // dedup only: the caller checked that it applies)
        const uint64_t chunk = level_chunk(n_front, lv_pp, ch_pp);
        level_buffers(chunk, lv_pp, ch_pp, a, d);
        if (chunk < n_front) {
            ch_cap = level_cap(n_front, ch_pp);
            ch_k.reserve(ch_cap);
            ch_c.reserve(ch_cap);
            a.ch_key = ch_k.p;
            a.ch_cnt = ch_c.p;
            d.ch_key = ch_k.p;
            d.ch_cnt = ch_c.p;
            a.ch_cap = d.ch_cap = ch_cap;
        }
        // the partitioned dedup writes the next frontier itself (as many
        // states as the table would hold)
        last_part = part_dedup_fits(ch_cap, n_front, tab_m, ch_pp * (double)n_front);
        if (chunk < n_front && !last_part) {
            h32[C32_OVERFLOW] = h32[C32_SPILL] = 0;
            return false;
        }
        nk.reserve(tab_m);
        nc.reserve(tab_m);
        PartArgs pa{};
        if (last_part) part_args(d, ch_cap, nk.p, nc.p, tab_m, pa);
        else blk.reserve(2 * COMPACT_MAX_BLOCKS);
        HIPCHK(hipEventRecord(ev0, st));
        HIPCHK(hipMemcpyAsync(result_save.p, result.p, (size_t)S * 8, hipMemcpyDeviceToDevice, st));
        if (!last_part) launch_fill_table(tk.p, tk.p + 1, tab_m, st);
        launch_reset_counters(c32.p, c64.p, st);
        queue_terminal();
        const uint32_t nch = queue_levels(a, d, 0, chunk, prune, true, last_part ? &pa : nullptr);
        HIPCHK(hipEventRecord(evc0, st));
        if (!last_part)
            launch_compact(tk.p, tk.p + 1, tab_m, nk.p, nc.p, tab_m, c32.p, blk.p, nullptr, st);
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipMemcpyAsync(pin32, c32.p, sizeof(uint32_t) * C32_N, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(pin64, c64.p, sizeof(unsigned long long) * C64_N,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipGetLastError());
        sync_host();
        std::memcpy(h32, pin32, sizeof(uint32_t) * C32_N);
        std::memcpy(h64, pin64, sizeof(unsigned long long) * C64_N);
        if (last_part && h32[C32_POVF] == 2u && !h32[C32_XOVF]) {
            // the unique states outgrew the next frontier: the buckets are
            // intact, so only the merge runs again, into buffers of the size
            // it counted
            nk.reserve(h64[C64_NEW]);
            nc.reserve(h64[C64_NEW]);
            pa.out_key = nk.p;
            pa.out_cnt = nc.p;
            pa.out_cap = nk.cap;
            launch_part_merge_again(pa, lib->n_cu, st);
            HIPCHK(hipEventRecord(ev1, st));
            HIPCHK(hipMemcpyAsync(pin32, c32.p, sizeof(uint32_t) * C32_N, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(pin64, c64.p, sizeof(unsigned long long) * C64_N,
                                  hipMemcpyDeviceToHost, st));
            HIPCHK(hipGetLastError());
            sync_host();
            std::memcpy(h32, pin32, sizeof(uint32_t) * C32_N);
            std::memcpy(h64, pin64, sizeof(unsigned long long) * C64_N);
            stats.part_remerges++;
        }
        level_times(nch);
        stats.fused_steps++;
        stats.donation_full += h32[C32_DONFULL];
        stats.extend_bytes += 16ull * n_front + 16ull * h64[C64_CHILDREN];
        stats.dedup_bytes += 16ull * (h64[C64_CHILDREN] + (level_prune ? 0 : h64[C64_PRUNED])) +
                             16ull * h64[C64_NEW];
        const bool pfail = last_part && (h32[C32_POVF] || (uint64_t)h32[C32_OUT] != h64[C64_NEW]);
        if (trace)
            fprintf(stderr,
                    "fused step %u: parents %llu, table %llu, caps %llu / %llu, children %llu, "
                    "new %llu, level overflow %u, table overflow %u, spilled %u, partitioned %d "
                    "(failed %d: flags %u, written %u, kept %llu, count sums %llu / %llu)\n",
                    g, (unsigned long long)n_front, (unsigned long long)tab_m,
                    (unsigned long long)lv_cap, (unsigned long long)ch_cap, h64[C64_CHW],
                    h64[C64_NEW], h32[C32_XOVF], h32[C32_OVERFLOW], h32[C32_SPILL], (int)last_part,
                    (int)pfail, h32[C32_POVF], h32[C32_OUT], h64[C64_CHILDREN], h64[C64_PSUM_IN],
                    h64[C64_PSUM_OUT]);
        if (h32[C32_XOVF] || h32[C32_OVERFLOW] || h32[C32_SPILL] || pfail) {
            copy(result.p, result_save.p, (size_t)S * 8, hipMemcpyDeviceToDevice);
            stats.fused_fallbacks++;
            if (pfail) stats.part_fallbacks++;
            last_part = false;
            return false;
        }
        if (last_part) {
            // the merge keeps the sum of the tiling counts it was given
            if (h64[C64_PSUM_IN] != h64[C64_PSUM_OUT])
                throw Error(D4_EDEVICE,
                            "internal error: partitioned dedup changed the count sum at step " +
                                std::to_string(g) + " (" + std::to_string(h64[C64_PSUM_IN]) +
                                " in, " + std::to_string(h64[C64_PSUM_OUT]) + " out, " +
                                std::to_string(h64[C64_NEW]) + " states, " +
                                std::to_string(h64[C64_CHW]) + " children written)");
            stats.part_steps++;
        }
        // (the level counters are the last chunk's)
        level_sizes(h64, n_front - (uint64_t)(nch - 1) * chunk, n_front);
        return true;
    }

    // One row step over the whole frontier.  With a memory cap (count mode)
    // it returns false, with nothing changed, when it would not fit.
    bool row_step(bool may_split = false) {
        if (n_front > 0xffffffffull) throw Error(D4_ERANGE, "frontier exceeds 2^32 records");
        const bool en = mode == D4_MODE_ENUM;
        const bool bounded = may_split && mem_cap && !en && n_front >= 2;
        const bool prune = !(flags & D4_FLAG_NO_PRUNE);
        // Table size: unique children <= children.  A small frontier gets an
        // exact bound (row-0 completions per record) or, failing that, an
        // exact count from a search-only pass; a large one is estimated from
        // the previous step's children per record, with a rerun at 4x the
        // size if the table still overflows.
        // (The row-0 bound ignores the lower rows and the piece counts, so
        // it can exceed the children many times over; past the first rows
        // the growth estimate is the tighter of the two.)
        uint64_t est_c = 0, est_u = 0;
        bool exact = false;
        const uint64_t g_c =
            (uint64_t)(n_front * std::min(std::max(grow_children, 1.0), 64.0) * 1.5) + 4096;
        const uint64_t g_u =
            std::min<uint64_t>(g_c, (uint64_t)(n_front * std::max(grow_unique, 0.25) * 1.5) + 4096);
        // (count mode past its first rows sizes from the growth estimate, as
        // the spill and table growth catch a short table: the bound, one more
        // host round trip, is taken only where new roots make growth unknown)
        pp_bound = 0.0;
        lib->seen_sizes(stats.rows_done, seen_lv, seen_ch);
        if (n_front <= (1u << 20) && (en || stats.rows_done < 2 || new_roots)) {
            const uint64_t bound = next_size();
            if (n_front && bound <= (1ull << 26)) pp_bound = (double)bound / (double)n_front;
            if (bound <= (1ull << 26)) {
                est_c = est_u = bound;
                exact = true;
                if (stats.rows_done >= 2 && g_u < bound) est_u = g_u;
            } else if (n_front <= kExactCountMax && en && (stats.rows_done < 2 || new_roots)) {
                // enum mode needs every child's slot: the first rows and new
                // roots (thousands of children per record) are counted
                // exactly; later rows take the growth estimate below, and a
                // table or edge buffer that overflows reruns the step
                est_c = est_u = count_children();
                stats.count_passes++;
                exact = true;
            } else if (n_front <= kExactCountMax && stats.rows_done < 2) {
                // count mode: a short table spills and grows, so the first
                // rows are sized from a generous per-record allowance rather
                // than by searching the row twice
                est_c = est_u = std::min<uint64_t>(bound, n_front * kFirstRowFanout + 65536);
                // (earlier searches on the library bound this row's children
                // per parent, which bound its unique states)
                if (seen_ch > 0)
                    est_c = est_u = std::min<uint64_t>(
                        est_u, (uint64_t)(1.2 * seen_ch * (double)n_front) + 65536);
                exact = true;
            }
        }
        if (!exact) {
            est_c = g_c;
            est_u = g_u;
        }
        // Count mode spills children that find no slot and grows the table
        // afterwards, so it is sized to the growth estimate without padding
        // (load 35-70% when the estimate holds); enum mode, which needs every
        // child's slot, keeps the padded size and reruns on overflow.
        uint64_t tab_m;
        if (en) {
            tab_m = round_slots(2 * est_u);
        } else {
            uint64_t t = est_u;
            if (!exact) t = std::min<uint64_t>(g_c, (uint64_t)(n_front * grow_unique_raw * 1.1) + 4096);
            tab_m = round_slots((uint64_t)(1.4 * table_scale * (double)t));
            if (!spill.ext) spill.reserve(2 * kSpillCap);
        }
        uint64_t est_e = en ? est_c + 64 : 0;
        if (en && logs.empty()) log0 = g;
        result_save.reserve(S);
        uint32_t h32[C32_N];
        unsigned long long h64[C64_N];
        bool trend = false;
        bool fused_ok = false, fused_skip = false;
        for (;;) {
            if (bounded && (tab_m > (1ull << 34) ||
                            16ull * tab_m + 16ull * std::min(est_u, tab_m) + 16ull * n_front +
                                    level_bytes(level_kernel(en, false) ? n_front : 0) +
                                    parked_bytes > mem_cap)) {
                // Growth slows from row to row, and a split frontier cannot
                // merge its parts' equal states (each part of a split at the
                // peak rows of 8x15 pentominoes x2 rebuilt nearly the whole
                // next frontier).  Before splitting, size the step from the
                // unique ratio extrapolated from the last two (their quotient
                // applied once more); if the children outgrow it, the spill
                // and table growth or a rerun catch it and the split follows.
                if (!exact && !trend && prev_unique_raw > grow_unique_raw) {
                    trend = true;
                    const double r = grow_unique_raw * grow_unique_raw / prev_unique_raw;
                    const uint64_t m2 =
                        round_slots((uint64_t)(1.4 * ((double)n_front * r * 1.1 + 4096)));
                    if (m2 < tab_m) {
                        tab_m = m2;
                        est_u = std::min<uint64_t>(est_u, (uint64_t)(n_front * r * 1.5) + 4096);
                        continue;
                    }
                }
                return false;
            }
            if (tab_m > (1ull << 34)) throw Error(D4_ERANGE, "dedup table exceeds 2^34 slots");
            // the allocation-free search always uses its whole table slice
            if (tk.ext) tab_m = tk.cap / 2;
            try {
                tk.reserve(2 * tab_m);
            } catch (const Error &e) {
                if (e.code != D4_ENOMEM || !bounded) throw;
                return false;
            }
            if (en) {
                if (est_e > 0xffffffffull) throw Error(D4_ERANGE, "edge count exceeds 2^32");
                e_parent.reserve(est_e);
                e_desc.reserve(est_e);
                e_slot.reserve(est_e);
            }
            if (!fused_skip && !en && lib->step_mode != D4_STEP_GENERAL && !noalloc &&
                n_front &&
                level_kernel(en, false) && tab_m <= kFusedMaxSlots &&
                (level_chunk(n_front, lv_pp0(), ch_pp0()) >= n_front || fused_chunks_ok(tab_m)) &&
                (!mem_cap || 48ull * tab_m + 16ull * n_front + level_bytes(n_front) +
                                     parked_bytes <= mem_cap)) {
                if (lib->step_mode == D4_STEP_GRAPH && !fused_skip && !g_off_once) {
                    // one replayed step; where replay does not apply, the fused chain
                    if (graph_rows(1, tab_m, prune)) return true;
                    if (g_fail) {
                        // the general path reruns it, with a larger table if that was short
                        fused_skip = true;
                        if (g_fail & 6) {
                            stats.retries++;
                            tab_m *= 4;
                        }
                        continue;
                    }
                }
                bool done = false;
                try {
                    done = fused_step(tab_m, prune, h32, h64);
                } catch (const Error &e) {
                    // no room for next-frontier buffers sized to the table:
                    // the general path sizes them to the unique count
                    if (e.code != D4_ENOMEM) throw;
                    h32[C32_OVERFLOW] = h32[C32_SPILL] = 0;
                }
                if (done) {
                    fused_ok = true;
                    break;
                }
                // a table overflow, a spill or a level buffer overflow: the
                // general path below reruns the step (a bigger table for the
                // first two, chunk reruns for the last)
                fused_skip = true;
                if (h32[C32_OVERFLOW] || h32[C32_SPILL]) {
                    stats.retries++;
                    tab_m = h32[C32_OVERFLOW]
                                ? 4 * tab_m
                                : round_slots(std::max<uint64_t>(
                                      2 * tab_m, 2 * ((uint64_t)h64[C64_NEW] + h32[C32_SPILL])));
                    continue;  // (the table is reserved again at its new size)
                }
            }
            HIPCHK(hipEventRecord(ev0, st));
            // terminal counts are added during the step: keep a copy so a
            // rerun after an overflow starts from the same counts
            HIPCHK(hipMemcpyAsync(result_save.p, result.p, (size_t)S * 8, hipMemcpyDeviceToDevice,
                                  st));
            launch_fill_table(tk.p, tk.p + 1, tab_m, st);
            launch_reset_counters(c32.p, c64.p, st);
            HIPCHK(hipEventRecord(evx0, st));
            launch_ranges(tab_m, (uint32_t)est_e, en, prune, false);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(evx1, st));
            HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
            sync_host();
            const bool lv = level_kernel(en, false);
            if (n_front && !lv) {
                float xms = 0;
                HIPCHK(hipEventElapsedTime(&xms, evx0, evx1));
                stats.extend_ms += xms;
                stats.extend_launches++;
            }
            stats.donation_full += h32[C32_DONFULL];  // (every attempt, reruns included)
            if (n_front) {
                stats.extend_bytes += 16ull * n_front + 16ull * h64[C64_CHILDREN];
                // level kernels: the dedup reads every child written (the
                // kept ones, and the pruned ones unless the level kernels
                // pruned them) and writes the new states
                if (lv)
                    stats.dedup_bytes +=
                        16ull * (h64[C64_CHILDREN] + (level_prune ? 0 : h64[C64_PRUNED])) +
                        16ull * h64[C64_NEW];
            }
            bool ovf = h32[C32_OVERFLOW] != 0;
            bool eovf = en && h32[C32_EDGE_OVF] != 0;
            // k_expand's spill buffer was full: rerun with four times the room
            // (the level kernels rerun their own chunks)
            const bool xovf = !ovf && !lv && h32[C32_XOVF] != 0;
            if (!ovf && !en && h32[C32_SPILL] && tk.ext)
                throw Error(D4_ERANGE, "allocation-free search: children spilled past the "
                                       "scratch's dedup table");
            if (!ovf && !en && h32[C32_SPILL]) {
                // move the table into one large enough for it and the spill
                const uint32_t ns = h32[C32_SPILL];
                const uint64_t m2 =
                    round_slots(std::max<uint64_t>(2 * tab_m, 2 * ((uint64_t)h64[C64_NEW] + ns)));
                bool room = !(mem_cap && 16ull * (tab_m + m2) + 16ull * n_front + parked_bytes >
                                             mem_cap);
                if (room) {
                    try {
                        tk2.reserve(2 * m2);
                    } catch (const Error &e) {
                        if (e.code != D4_ENOMEM) throw;
                        room = false;
                    }
                }
                if (room) {
                    launch_rehash(tk.p, tab_m, tk2.p, m2, spill.p, ns, c32.p, c64.p, st);
                    HIPCHK(hipGetLastError());
                    HIPCHK(hipMemcpyAsync(h32, c32.p, sizeof h32, hipMemcpyDeviceToHost, st));
                    HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
                    sync_host();
                    ovf = h32[C32_OVERFLOW] != 0;
                    if (!ovf) {
                        tk.swap(tk2);
                        // the old table goes back to the allocator, where
                        // the next frontier can take its memory
                        tk2.release();
                        tab_m = m2;
                        stats.rehashes++;
                    }
                } else {
                    ovf = true;
                }
            }
            if (!ovf && !eovf && !xovf) break;
            // the step is a pure function of the frontier: rerun it bigger
            stats.retries++;
            copy(result.p, result_save.p, (size_t)S * 8, hipMemcpyDeviceToDevice);
            if (ovf && tk.ext)
                throw Error(D4_ERANGE, "allocation-free search: the row's states exceed the "
                                       "scratch's dedup table");
            if (ovf) tab_m *= 4;
            if (xovf) xp_cap *= 4;
            if (eovf) est_e = (uint64_t)h32[C32_EDGES] + 1024;
        }
        const uint64_t U = h64[C64_NEW];
        check_row_states(U);
        if (!fused_ok) {
            try {
                nk.reserve(U);
                nc.reserve(U);
            } catch (const Error &e) {
                if (e.code != D4_ENOMEM || !bounded) throw;
                copy(result.p, result_save.p, (size_t)S * 8, hipMemcpyDeviceToDevice);
                return false;
            }
        }
        stats.children += h64[C64_CHILDREN];
        stats.pruned += h64[C64_PRUNED];
        stats.donated += h64[C64_DONATED];
        stats.records_extended += n_front;
        stats.unique += U;
        if (n_front) {
            prev_unique_raw = stats.rows_done ? grow_unique_raw : 0.0;
            grow_unique = std::max(1.0, (double)U / (double)n_front);
            grow_unique_raw = std::max(0.01, (double)U / (double)n_front);
            grow_children = std::max(1.0, (double)h64[C64_CHILDREN] / (double)n_front);
        }
        uint32_t n_out = 0;
        const uint32_t n_edges = en ? h32[C32_EDGES] : 0;
        if (fused_ok) {
            n_out = h32[C32_OUT];  // (compacted inside the fused chain)
        } else {
            if (en) slot_uid.reserve(tab_m);
            HIPCHK(hipEventRecord(evc0, st));
            blk.reserve(2 * COMPACT_MAX_BLOCKS);
            launch_compact(tk.p, tk.p + 1, tab_m, nk.p, nc.p, U, c32.p, blk.p,
                           en ? slot_uid.p : nullptr, st);
            HIPCHK(hipMemcpyAsync(&n_out, c32.p + C32_OUT, 4, hipMemcpyDeviceToHost, st));
            if (en && n_edges) {
                e_child.reserve(n_edges);
                launch_remap_edges(e_slot.p, slot_uid.p, e_child.p, n_edges, st);
            }
            HIPCHK(hipEventRecord(ev1, st));
            HIPCHK(hipGetLastError());
            sync_host();
        }
        check_compacted(n_out, U);
        float ms = 0, cms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&cms, evc0, ev1));
        stats.gpu_ms += ms;
        stats.compact_ms += cms;
        if (!(fused_ok && last_part)) {
            stats.compact_launches++;
            stats.compact_bytes += 8ull * tab_m + 24ull * U;
        }
        if (en) {
            // the step's frontier and edges stay on the device as its log;
            // the next step gets fresh buffers from the allocator cache
            logs.emplace_back();
            DevLog &lg = logs.back();
            lg.keys.swap(fk);
            lg.n_keys = n_front;
            lg.par.swap(e_parent);
            lg.chi.swap(e_child);
            lg.desc.swap(e_desc);
            lg.n_edges = n_edges;
        }
        fk.swap(nk);
        fc.swap(nc);
        n_front = U;
        if (!U) {
            fk.reserve(1);
            fc.reserve(1);
        }
        stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, U);
        stats.rows_done++;
        new_roots = 0;
        return true;
    }

    // One row step whose dedup table is split into P partitions by key hash
    // (low bits of the hash; the home slot takes the high ones): pass p
    // reads every parent, in chunks (host parts copied in), and keeps only
    // the children of partition p, so each pass's table holds about a P-th
    // of the unique states while equal states still meet in one table.  Each
    // partition is compacted into a part of the next frontier, in device
    // memory or, with a host budget, in pinned host memory.  Unlike parking
    // halves of the frontier, nothing is searched twice.  Returns false,
    // with nothing changed, when no partition count fits.
    bool streamed_step(bool prune) {
        struct Src {
            const uint64_t *k, *c, *h;  // device keys / counts, or host [keys, counts]
            uint64_t n;
        };
        std::vector<Src> src;
        uint64_t dev_src = 0, max_host = 0;
        if (fparts.empty()) {
            src.push_back({fk.p, fc.p, nullptr, n_front});
            dev_src = 16 * n_front;
        } else {
            for (auto &q : fparts) {
                if (q->host) {
                    src.push_back({nullptr, nullptr, q->host, q->n});
                    max_host = std::max(max_host, q->n);
                } else {
                    src.push_back({q->k.p, q->c.p, nullptr, q->n});
                    dev_src += 16 * q->n;
                }
            }
        }
        // unique children: the last step's ratio, extrapolated once more
        // when growth speeds up
        double r = std::max(grow_unique_raw, 0.01);
        if (prev_unique_raw > 0 && r > prev_unique_raw) r = r * r / prev_unique_raw;
        const double est_u = (double)n_front * r * 1.1 + 4096.0;
        const uint64_t cap = mem_cap ? mem_cap : (uint64_t)(0.8 * (double)available_device_bytes());
        const uint64_t stage_n = std::min<uint64_t>(max_host, 1ull << 26);
        const uint64_t fixed = dev_src + parked_bytes + level_budget() + 16 * stage_n +
                               (host_cap ? 0 : (uint64_t)(16.0 * est_u));
        if (fixed >= cap) return false;
        const double avail = (double)(cap - fixed);
        // (every pass reads all parents: past 64 partitions, parking halves
        // of the frontier costs less)
        uint32_t P = 2;
        while (P < kMaxPartitions && 2.0 * 38.4 * est_u / P > avail) P *= 2;
        if (2.0 * 38.4 * est_u / P > avail) return false;
        if (host_cap && host_parked + (uint64_t)(16.0 * est_u) > host_cap + (host_cap >> 1))
            return false;
        sk.reserve(std::max<uint64_t>(stage_n, 1));
        sc.reserve(std::max<uint64_t>(stage_n, 1));
        result_save.reserve(S);
        if (!spill.ext) spill.reserve(2 * kSpillCap);
        HIPCHK(hipMemcpyAsync(result_save.p, result.p, (size_t)S * 8, hipMemcpyDeviceToDevice, st));
        std::vector<std::unique_ptr<Part>> out;
        uint64_t U = 0, children = 0, pruned = 0;
        for (uint32_t p = 0; p < P; p++) {
            for (uint32_t attempt = 0;; attempt++) {
                if (attempt > 8) throw Error(D4_ERANGE, "a streamed partition outgrows its table");
                const uint64_t tab_m = round_slots((uint64_t)(1.4 * table_scale * est_u / P) << attempt);
                tk.reserve(2 * tab_m);
                launch_fill_table(tk.p, tk.p + 1, tab_m, st);
                launch_reset_counters(c32.p, c64.p, st);
                bool ovf = false;
                for (const Src &x : src) {
                    if (!x.h) {
                        run_levels(x.k, x.c, x.n, tab_m, prune, P - 1, p, p == 0);
                    } else {
                        for (uint64_t b = 0; b < x.n && !ovf; b += stage_n) {
                            const uint64_t m = std::min(stage_n, x.n - b);
                            HIPCHK(hipMemcpyAsync(sk.p, x.h + b, 8 * m, hipMemcpyHostToDevice, st));
                            HIPCHK(hipMemcpyAsync(sc.p, x.h + x.n + b, 8 * m,
                                                  hipMemcpyHostToDevice, st));
                            run_levels(sk.p, sc.p, m, tab_m, prune, P - 1, p, p == 0);
                            uint32_t f = 0;
                            HIPCHK(hipMemcpyAsync(&f, c32.p + C32_OVERFLOW, 4, hipMemcpyDeviceToHost, st));
                            sync_host();
                            ovf = f != 0;
                        }
                    }
                    uint32_t h[C32_N];
                    HIPCHK(hipMemcpyAsync(h, c32.p, sizeof h, hipMemcpyDeviceToHost, st));
                    sync_host();
                    ovf = h[C32_OVERFLOW] || h[C32_SPILL];
                    if (ovf) break;
                }
                if (ovf) {
                    // a larger table; the first pass's terminal counts are undone
                    stats.retries++;
                    if (p == 0)
                        copy(result.p, result_save.p, (size_t)S * 8, hipMemcpyDeviceToDevice);
                    continue;
                }
                unsigned long long h64[C64_N];
                HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
                sync_host();
                const uint64_t u = h64[C64_NEW];
                check_row_states(U + u);
                nk.reserve(std::max<uint64_t>(u, 1));
                nc.reserve(std::max<uint64_t>(u, 1));
                blk.reserve(2 * COMPACT_MAX_BLOCKS);
                launch_compact(tk.p, tk.p + 1, tab_m, nk.p, nc.p, u, c32.p, blk.p, nullptr, st);
                uint32_t n_out = 0;
                HIPCHK(hipMemcpyAsync(&n_out, c32.p + C32_OUT, 4, hipMemcpyDeviceToHost, st));
                sync_host();
                check_compacted(n_out, u);
                auto q = std::make_unique<Part>();
                q->g = g + 1;
                q->n = u;
                if (host_cap && u) {
                    if (host_parked + 16 * u > host_cap)
                        throw Error(D4_ENOMEM, "frontier parts exceed the host memory budget");
                    HIPCHK(hipHostMalloc((void **)&q->host, 16 * u));
                    HIPCHK(hipMemcpyAsync(q->host, nk.p, 8 * u, hipMemcpyDeviceToHost, st));
                    HIPCHK(hipMemcpyAsync(q->host + u, nc.p, 8 * u, hipMemcpyDeviceToHost, st));
                    sync_host();
                    host_parked += 16 * u;
                    stats.host_parked_peak = std::max<uint64_t>(stats.host_parked_peak, host_parked);
                } else {
                    q->k.swap(nk);
                    q->c.swap(nc);
                }
                out.push_back(std::move(q));
                U += u;
                children += h64[C64_CHILDREN];
                pruned += h64[C64_PRUNED];
                stats.stream_passes++;
                if (trace)
                    fprintf(stderr, "streamed step %u: partition %u of %u, table %llu, new %llu\n",
                            g, p, P, (unsigned long long)tab_m, (unsigned long long)u);
                break;
            }
        }
        // the parents are done with; the partitions are the next frontier
        for (auto &q : fparts)
            if (q->host) host_parked -= 16 * q->n;
        fparts.clear();
        tk.release();
        stats.children += children;
        stats.pruned += pruned;
        stats.records_extended += n_front;
        stats.unique += U;
        if (n_front) {
            prev_unique_raw = stats.rows_done ? grow_unique_raw : 0.0;
            grow_unique = std::max(1.0, (double)U / (double)n_front);
            grow_unique_raw = std::max(0.01, (double)U / (double)n_front);
            grow_children = std::max(1.0, (double)children / (double)n_front);
        }
        stats.peak_frontier = std::max<uint64_t>(stats.peak_frontier, U);
        stats.rows_done++;
        stats.streamed_steps++;
        new_roots = 0;
        n_front = U;
        fparts = std::move(out);
        // gathered into fk / fc when the frontier fits with room to spare
        if (16 * U <= cap / 4) {
            fk.release();
            fc.release();
            fk.reserve(std::max<uint64_t>(U, 1));
            fc.reserve(std::max<uint64_t>(U, 1));
            uint64_t at = 0;
            for (auto &q : fparts) {
                if (q->host) {
                    HIPCHK(hipMemcpyAsync(fk.p + at, q->host, 8 * q->n, hipMemcpyHostToDevice, st));
                    HIPCHK(hipMemcpyAsync(fc.p + at, q->host + q->n, 8 * q->n,
                                          hipMemcpyHostToDevice, st));
                } else if (q->n) {
                    HIPCHK(hipMemcpyAsync(fk.p + at, q->k.p, 8 * q->n, hipMemcpyDeviceToDevice, st));
                    HIPCHK(hipMemcpyAsync(fc.p + at, q->c.p, 8 * q->n, hipMemcpyDeviceToDevice, st));
                }
                at += q->n;
            }
            sync_host();
            for (auto &q : fparts)
                if (q->host) host_parked -= 16 * q->n;
            fparts.clear();
        } else {
            fk.release();
            fc.release();
        }
        return true;
    }

    // The parts of a frontier left by a streamed step, parked (as halves
    // are) and the last one resumed: depth first from here on.
    void park_parts() {
        for (auto &q : fparts) {
            q->g = g;
            q->grow_unique = grow_unique;
            q->grow_unique_raw = grow_unique_raw;
            q->prev_unique_raw = prev_unique_raw;
            q->grow_children = grow_children;
            if (!q->host) parked_bytes += 16 * q->n;
            parked.push_back(std::move(q));
            stats.splits++;
        }
        fparts.clear();
        resume();
    }

    // Park the upper half of the frontier, to resume at this step later.
    void split() {
        // the failed step's table and next-frontier buffers hold no state:
        // return them to the allocator before the parked half is allocated
        tk.release();
        tk2.release();
        nk.release();
        nc.release();
        const uint64_t keep = n_front / 2, m = n_front - keep;
        auto p = std::make_unique<Part>();
        p->g = g;
        p->n = m;
        p->grow_unique = grow_unique;
        p->grow_unique_raw = grow_unique_raw;
        p->prev_unique_raw = prev_unique_raw;
        p->grow_children = grow_children;
        if (host_cap) {
            if (host_parked + 16 * m > host_cap)
                throw Error(D4_ENOMEM, "parked frontier parts exceed the host memory budget");
            HIPCHK(hipHostMalloc((void **)&p->host, 16 * m));
            HIPCHK(hipMemcpyAsync(p->host, fk.p + keep, 8 * m, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(p->host + m, fc.p + keep, 8 * m, hipMemcpyDeviceToHost, st));
            sync_host();
            host_parked += 16 * m;
            stats.host_parked_peak = std::max<uint64_t>(stats.host_parked_peak, host_parked);
        } else {
            p->k.reserve(m);
            p->c.reserve(m);
            HIPCHK(hipMemcpyAsync(p->k.p, fk.p + keep, 8 * m, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(p->c.p, fc.p + keep, 8 * m, hipMemcpyDeviceToDevice, st));
            sync_host();
            parked_bytes += 16 * m;
        }
        parked.push_back(std::move(p));
        new_roots = std::min<uint64_t>(new_roots, keep);
        n_front = keep;
        stats.splits++;
    }

    void resume() {
        std::unique_ptr<Part> p = std::move(parked.back());
        parked.pop_back();
        if (p->host) {
            fk.reserve(std::max<uint64_t>(p->n, 1));
            fc.reserve(std::max<uint64_t>(p->n, 1));
            HIPCHK(hipMemcpyAsync(fk.p, p->host, 8 * p->n, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(fc.p, p->host + p->n, 8 * p->n, hipMemcpyHostToDevice, st));
            sync_host();
            host_parked -= 16 * p->n;
        } else {
            parked_bytes -= 16 * p->n;
            fk.swap(p->k);
            fc.swap(p->c);
        }
        n_front = p->n;
        g = p->g;
        grow_unique = p->grow_unique;
        grow_unique_raw = p->grow_unique_raw;
        prev_unique_raw = p->prev_unique_raw;
        grow_children = p->grow_children;
        new_roots = 0;
    }

    // Count mode under a memory cap: every row of every instance, parking
    // frontier halves that do not fit and finishing them one after the
    // other (depth first); tiling counts add up per instance.
    void run_bounded() {
        completed.clear();
        upload_pending();
        uint32_t gend = g;
        for (uint32_t i = 0; i < S; i++)
            if (active[i]) gend = std::max(gend, start[i] + heights[i]);
        for (;;) {
            while (n_front && g <= gend) {
                const bool prune_ = !(flags & D4_FLAG_NO_PRUNE);
                if (!fparts.empty()) {
                    // a frontier left in parts streams through the next step
                    // too; where no partitioning fits, its parts are parked
                    // and finished one after the other
                    if (!streamed_step(prune_)) {
                        park_parts();
                        continue;
                    }
                    if (trace)
                        fprintf(stderr, "run step %u: frontier %llu in %zu parts\n", g,
                                (unsigned long long)n_front, fparts.size());
                    g++;
                    continue;
                }
                // past the first rows (whose growth is unknown), replay many
                // steps per host round trip; a failed step takes the path below
                if (stats.rows_done >= 2 && !new_roots) {
                    const uint32_t k = graph_rows(gend - g + 1, 0, !(flags & D4_FLAG_NO_PRUNE));
                    if (k) {
                        g += k;
                        if (trace)
                            fprintf(stderr, "run step %u: frontier %llu, %zu parts parked\n",
                                    g - 1, (unsigned long long)n_front, parked.size());
                        continue;
                    }
                    // (a step that failed is not replayed again)
                    g_off_once = g_fail != 0;
                }
                const bool stepped = row_step(true);
                g_off_once = false;
                if (stepped) {
                    g++;
                    take_sample();
                    g--;
                    if (trace)
                        fprintf(stderr, "run step %u: frontier %llu, %zu parts parked\n", g,
                                (unsigned long long)n_front, parked.size());
                    g++;
                    continue;
                }
                // the row does not fit: stream it through partitioned tables
                // (level kernels), else park half of the frontier
                if (stream_ok && level_kernel(false, false) && n_front >= 2 &&
                    streamed_step(prune_)) {
                    if (trace)
                        fprintf(stderr, "run step %u: frontier %llu in %zu parts (streamed)\n", g,
                                (unsigned long long)n_front, fparts.size());
                    g++;
                    continue;
                }
                if (n_front < 2) throw Error(D4_ENOMEM, "one state's row step exceeds the memory cap");
                split();
            }
            if (parked.empty()) break;
            resume();
        }
        g = gend + 1;
        n_front = 0;
        fparts.clear();
        host_parked = 0;
        copy(host_result.data(), result.p, (size_t)S * 8, hipMemcpyDeviceToHost);
        for (uint32_t i = 0; i < S; i++)
            if (active[i]) {
                completed.push_back(i);
                active[i] = 0;
                n_active--;
            }
    }

    // Extend the whole frontier.  Split levels spread each record over 64 /
    // 4096 work items for tiny frontiers; the roots admitted this step (a
    // contiguous range at the end) always get their own launch with two
    // split levels, because their first row has the largest subtrees and a
    // mixed frontier would otherwise leave each of them to a single lane.
    void launch_ranges(uint64_t tab_m, uint32_t e_cap, bool en, bool prune, bool count_only) {
        if (!n_front) return;
        if (level_kernel(en, count_only)) {
            run_levels(tab_m, prune);
            return;
        }
        const bool xp = expand_kernel(en, count_only);
        don_a.reserve(xp ? EXPAND_WORDS * xp_cap : 2u << 20);  // 1 Mi donated items per launch
        don_b.reserve(xp ? EXPAND_WORDS * xp_cap : 2u << 20);
        auto split_of = [](uint64_t n) {
            return n <= 1024 ? 2u : n <= (1u << 16) ? 1u : 0u;
        };
        const uint64_t roots = std::min<uint64_t>(new_roots, n_front);
        const uint64_t old = n_front - roots;
        if (xp) {
            auto go = [&](uint64_t b, uint64_t n, uint32_t split) {
                ExtendArgs a = args(tab_m, 0, b, n, split);
                a.don_cap = (uint32_t)std::min<uint64_t>(don_a.cap / EXPAND_WORDS, xp_cap);
                a.budget = lib->xp_budget;
                launch_expand(a, prune, lib->ext_threads, lib->n_cu, st);
            };
            if (old && roots) {
                go(0, old, split_of(old));
                launch_reset_work(c64.p, st);
                go(old, roots, 2);
            } else {
                go(0, n_front, split_of(n_front));
            }
        } else if (old && roots) {
            launch_extend(args(tab_m, e_cap, 0, old, split_of(old)), en, prune, count_only,
                          lib->ext_threads, lib->n_cu, st);
            launch_reset_work(c64.p, st);
            launch_extend(args(tab_m, e_cap, old, roots, 2), en, prune, count_only,
                          lib->ext_threads, lib->n_cu, st);
        } else {
            launch_extend(args(tab_m, e_cap, 0, n_front, split_of(n_front)), en, prune,
                          count_only, lib->ext_threads, lib->n_cu, st);
        }
        run_donations(tab_m, e_cap, en, prune, count_only);
    }

    // Children the current frontier will emit (after pruning): the extend
    // search with inserts switched off.
    uint64_t count_children() {
        if (!n_front) return 0;
        const bool prune = !(flags & D4_FLAG_NO_PRUNE);
        launch_reset_counters(c32.p, c64.p, st);
        launch_ranges(1, 0, false, prune, true);
        HIPCHK(hipGetLastError());
        unsigned long long h64[C64_N];
        HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
        sync_host();
        return h64[C64_CHILDREN];
    }

    // Sample of the states of a single instance's frontier at a given
    // height (rows left), taken as the search passes it (bounded runs
    // included): the first states of the frontier, up to sample_cap.
    uint32_t sample_h = 0;
    uint64_t sample_cap = 0;
    std::vector<uint64_t> sample_keys, sample_cnt;
    void take_sample() {
        if (!fparts.empty()) return;  // (a frontier in parts is not sampled)
        if (!sample_h || S != 1 || !active[0] || sample_keys.size() >= sample_cap) return;
        if (heights[0] - std::min(heights[0], g - start[0]) != sample_h) return;
        const uint64_t k = std::min<uint64_t>(sample_cap - sample_keys.size(), n_front);
        if (!k) return;
        const size_t at = sample_keys.size();
        sample_keys.resize(at + k);
        sample_cnt.resize(at + k);
        copy(sample_keys.data() + at, fk.p, 8 * k, hipMemcpyDeviceToHost);
        copy(sample_cnt.data() + at, fc.p, 8 * k, hipMemcpyDeviceToHost);
    }

    // Per-bucket sizes of the current frontier (bucket = the current row's
    // open cells), cached until the frontier changes.
    unsigned long long bucket_n[256] = {}, bucket_next[256] = {};
    uint64_t bucket_g = ~0ull;
    DevBuf<unsigned long long> bucket_d;
    void buckets() {
        upload_pending();
        if (bucket_g == ((uint64_t)g << 32 | stats.rows_done)) return;
        bucket_d.reserve(512);
        if (n_front)
            launch_buckets(fk.p, n_front, g, dstart.p, bw.p, bw_stride, hgt.p, inst_shift,
                           occ_mask, lib->d_comp, bucket_d.p, lib->n_cu, st);
        else
            launch_zero_u32(reinterpret_cast<uint32_t *>(bucket_d.p), 2 * 512, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(bucket_n, bucket_d.p, sizeof bucket_n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(bucket_next, bucket_d.p + 256, sizeof bucket_next,
                              hipMemcpyDeviceToHost, st));
        sync_host();
        bucket_g = (uint64_t)g << 32 | stats.rows_done;
    }

    uint64_t next_size() {
        upload_pending();
        if (!n_front) return 0;
        launch_reset_counters(c32.p, c64.p, st);
        launch_next_size(fk.p, (uint32_t)n_front, g, dstart.p, bw.p, bw_stride, hgt.p, inst_shift,
                         occ_mask, lib->d_comp, c32.p, c64.p, lib->n_cu, st);
        unsigned long long h64[C64_N];
        HIPCHK(hipMemcpyAsync(h64, c64.p, sizeof h64, hipMemcpyDeviceToHost, st));
        sync_host();
        return h64[C64_NEXTSIZE];
    }

    // Drop enum logs no unfinished instance can need any more.
    void trim_logs() {
        if (keep_logs || logs.empty()) return;
        uint32_t keep_from = g;
        for (uint32_t i = 0; i < S; i++)
            if (active[i]) keep_from = std::min(keep_from, start[i]);
        while (!logs.empty() && log0 < keep_from) {
            logs.pop_front();
            log0++;
        }
    }

    // ---- reconstruction (enum mode) --------------------------------------

    // Each solution of the instance in `slot` (finished at step
    // start + height) as its sequence of candidate ids in placement order:
    // the back-pointers are walked from its terminal states down to its root.
    std::vector<std::vector<uint8_t>> solutions_of(uint32_t slot, bool single) const {
        if (mode != D4_MODE_ENUM) throw Error(D4_ESTATE, "solutions need D4_MODE_ENUM");
        if (slot >= S) throw Error(D4_EINVAL, "instance out of range");
        if (active[slot]) throw Error(D4_ESTATE, "search not finished");
        const uint32_t H = heights[slot], s0 = start[slot];
        if (s0 < log0 || s0 + H >= log0 + logs.size())
            throw Error(D4_ESTATE, "solutions of this instance are no longer held");
        std::vector<RowLog> hl(H + 1);
        for (uint32_t r = 0; r <= H; r++) host_log(s0 + r - log0, hl[r], r < H);
        auto L = [&](uint32_t r) -> const RowLog & { return hl[r]; };
        auto inst_of = [&](uint64_t key) -> uint32_t {
            return inst_shift < 64 ? (uint32_t)(key >> inst_shift) : 0u;
        };
        // reverse adjacency per row: child -> edges
        std::vector<std::vector<uint32_t>> first(H), order(H);
        for (uint32_t r = 0; r < H; r++) {
            const RowLog &lg = L(r);
            size_t nchild = L(r + 1).keys.size();
            std::vector<uint32_t> cnt(nchild + 1, 0);
            for (uint32_t c : lg.e_child) cnt[c + 1]++;
            for (size_t j = 1; j <= nchild; j++) cnt[j] += cnt[j - 1];
            std::vector<uint32_t> ord(lg.e_child.size());
            std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
            for (uint32_t e = 0; e < lg.e_child.size(); e++) ord[pos[lg.e_child[e]]++] = e;
            first[r] = std::move(cnt);
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
            const RowLog &lg = L(r - 1);
            for (uint32_t j = first[r - 1][u]; j < first[r - 1][u + 1]; j++) {
                uint32_t e = order[r - 1][j];
                if (inst_of(L(r - 1).keys[lg.e_parent[e]]) != slot) continue;
                path[r - 1] = lg.e_desc[e];
                walk(r - 1, lg.e_parent[e]);
            }
        };
        const std::vector<uint64_t> &term = L(H).keys;
        for (uint32_t u = 0; u < term.size(); u++)
            if (inst_of(term[u]) == slot) walk(H, u);
        std::vector<std::vector<uint8_t>> out;
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

    void host_log(size_t j, RowLog &out, bool edges) const {
        const DevLog &lg = logs[j];
        out.keys.resize(lg.n_keys);
        if (lg.n_keys)
            copy(out.keys.data(), lg.keys.p, lg.n_keys * 8, hipMemcpyDeviceToHost);
        if (!edges) return;
        out.e_parent.resize(lg.n_edges);
        out.e_child.resize(lg.n_edges);
        out.e_desc.resize(lg.n_edges);
        if (lg.n_edges) {
            copy(out.e_parent.data(), lg.par.p, lg.n_edges * 4, hipMemcpyDeviceToHost);
            copy(out.e_child.data(), lg.chi.p, lg.n_edges * 4, hipMemcpyDeviceToHost);
            copy(out.e_desc.data(), lg.desc.p, lg.n_edges * 8, hipMemcpyDeviceToHost);
        }
    }

    // First solutions in the CPU searcher's order (solve(..., true)) of
    // finished instances, without listing the others, on the device.  Per
    // group of instances sharing a start step: states that reach one of the
    // group's terminal states are marked backwards; in the same pass every
    // marked state keeps its smallest completion (candidate ids compared in
    // placement order) to a marked child; then one thread per instance walks
    // from its root along those.  Two completions of one state never have
    // one as a prefix of the other, and smaller completions order the whole
    // sequences, so the walk gives the lexicographic minimum.  An instance
    // with no tiling gets an empty list.  Cost: one pass over the group's
    // edges, whatever the number of instances or solutions.
    std::map<uint32_t, std::vector<std::vector<uint8_t>>> first_solutions(
        const std::vector<uint32_t> &slots) {
        if (mode != D4_MODE_ENUM) throw Error(D4_ESTATE, "solutions need D4_MODE_ENUM");
        std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> groups;
        for (uint32_t slot : slots) {
            if (slot >= S) throw Error(D4_EINVAL, "instance out of range");
            if (active[slot]) throw Error(D4_ESTATE, "search not finished");
            if (start[slot] < log0 || start[slot] + heights[slot] >= log0 + logs.size())
                throw Error(D4_ESTATE, "solutions of this instance are no longer held");
            groups[{start[slot], heights[slot]}].push_back(slot);
        }
        std::map<uint32_t, std::vector<std::vector<uint8_t>>> out;
        for (auto &gr : groups) {
            const uint32_t s0 = gr.first.first, H = gr.first.second;
            const std::vector<uint32_t> &mem = gr.second;
            std::vector<uint8_t> member(S, 0);
            for (uint32_t slot : mem) member[slot] = 1;
            DevBuf<uint8_t> d_member;
            DevBuf<uint32_t> d_root, d_slots;
            d_member.reserve(S);
            d_root.reserve(S);
            d_slots.reserve(mem.size());
            HIPCHK(hipMemcpyAsync(d_member.p, member.data(), S, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemsetAsync(d_root.p, 0xff, (size_t)S * 4, st));
            HIPCHK(hipMemcpyAsync(d_slots.p, mem.data(), mem.size() * 4, hipMemcpyHostToDevice, st));
            std::vector<DevBuf<uint8_t>> live(H + 1);
            std::vector<DevBuf<unsigned long long>> best(H);
            std::vector<DevBuf<uint32_t>> next(H);
            auto LG = [&](uint32_t r) -> const DevLog & { return logs[s0 + r - log0]; };
            for (uint32_t r = 0; r <= H; r++) {
                live[r].reserve(std::max<uint64_t>(LG(r).n_keys, 1));
                HIPCHK(hipMemsetAsync(live[r].p, 0, std::max<uint64_t>(LG(r).n_keys, 1), st));
            }
            launch_first_solutions_ends(LG(H).keys.p, LG(H).n_keys, LG(0).keys.p, LG(0).n_keys,
                                        inst_shift, d_member.p, live[H].p, d_root.p, st);
            for (uint32_t r = H; r-- > 0;) {
                const DevLog &lg = LG(r);
                const uint64_t nk = std::max<uint64_t>(lg.n_keys, 1);
                best[r].reserve(nk);
                next[r].reserve(nk);
                HIPCHK(hipMemsetAsync(best[r].p, 0xff, nk * 8, st));
                HIPCHK(hipMemsetAsync(next[r].p, 0xff, nk * 4, st));
                if (lg.n_edges)
                    launch_first_solutions_row(lg.par.p, lg.chi.p, lg.desc.p, lg.n_edges,
                                               live[r + 1].p, live[r].p, best[r].p, next[r].p, st);
            }
            std::vector<const unsigned long long *> hb(std::max(H, 1u), nullptr);
            std::vector<const uint32_t *> hn(std::max(H, 1u), nullptr);
            for (uint32_t r = 0; r < H; r++) {
                hb[r] = best[r].p;
                hn[r] = next[r].p;
            }
            DevBuf<uint64_t> d_ptr, d_out;
            d_ptr.reserve(2 * hb.size());
            d_out.reserve(mem.size() * (H + 1));
            HIPCHK(hipMemcpyAsync(d_ptr.p, hb.data(), hb.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d_ptr.p + hb.size(), hn.data(), hn.size() * 8,
                                  hipMemcpyHostToDevice, st));
            launch_first_solutions_walk(d_slots.p, (uint32_t)mem.size(), d_root.p, live[0].p,
                                        reinterpret_cast<const unsigned long long *const *>(d_ptr.p),
                                        reinterpret_cast<const uint32_t *const *>(d_ptr.p + hb.size()),
                                        H, d_out.p, st);
            HIPCHK(hipGetLastError());
            std::vector<uint64_t> ho(mem.size() * (H + 1));
            HIPCHK(hipMemcpyAsync(ho.data(), d_out.p, ho.size() * 8, hipMemcpyDeviceToHost, st));
            sync_host();
            for (size_t j = 0; j < mem.size(); j++) {
                const uint64_t *o = &ho[j * (H + 1)];
                if (!o[H]) {
                    out[mem[j]] = {};
                    continue;
                }
                std::vector<uint8_t> seq;
                for (uint32_t r = 0; r < H; r++)
                    for (int b = 0; b < 8 && ((o[r] >> (8 * b)) & 0xffu); b++)
                        seq.push_back((uint8_t)(((o[r] >> (8 * b)) & 0xffu) - 1));
                out[mem[j]] = {seq};
            }
        }
        return out;
    }

    std::vector<std::vector<uint8_t>> first_solution_of(uint32_t slot) {
        return first_solutions({slot})[slot];
    }

    // Steps {piece_id, trs_id, y, x} of a candidate sequence, replayed on the
    // board (anchor = lowest open cell, Piece.inl:5-24).
    void encode(uint32_t slot, const std::vector<uint8_t> &seq, std::vector<uint8_t> &buf) const {
        uint8_t open[MAX_ROWS + 8];
        std::memcpy(open, &boards[(size_t)slot * (MAX_ROWS + 8)], sizeof open);
        buf.push_back((uint8_t)seq.size());
        uint32_t r = 0;
        for (uint8_t k : seq) {
            while (r < heights[slot] && !open[r]) r++;
            if (r >= heights[slot]) throw Error(D4_EDEVICE, "inconsistent back-pointers");
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

    std::vector<uint8_t> solution_bytes(uint32_t slot, bool single, uint64_t *nsol) {
        auto seqs = single ? first_solution_of(slot) : solutions_of(slot, false);
        std::vector<uint8_t> buf;
        for (auto &sq : seqs) encode(slot, sq, buf);
        if (nsol) *nsol = seqs.size();
        return buf;
    }

    void solutions(uint32_t slot, int single, uint8_t *out, size_t cap, size_t *len,
                   uint64_t *nsol) {
        auto buf = solution_bytes(slot, single != 0, nsol);
        if (len) *len = buf.size();
        if (out && buf.size() <= cap) std::memcpy(out, buf.data(), buf.size());
        else if (out && cap) throw Error(D4_ERANGE, "solution buffer too small");
    }
};

// Engine with every instance admitted at step 0 (batch and single search).
static std::unique_ptr<Engine> make_batch(const d4_library *lib, uint32_t n, const uint8_t *b,
                                          uint32_t stride, const uint32_t *h, uint32_t mode,
                                          uint32_t flags) {
    if (!n) throw Error(D4_EINVAL, "no instances");
    uint32_t hm = 0;
    for (uint32_t i = 0; i < n; i++) hm = std::max(hm, h[i]);
    auto e = std::make_unique<Engine>(lib, n, hm, mode, flags, true);
    for (uint32_t i = 0; i < n; i++) e->admit(i, b + (size_t)i * stride, h[i]);
    e->upload_pending();
    if (mode == D4_MODE_COUNT) e->mem_cap = (uint64_t)(0.8 * (double)available_device_bytes());
    return e;
}

}  // namespace

struct d4_searcher {
    std::unique_ptr<Engine> e;
};

// Request loop with continuous batching: "solve all" requests run in a
// count-mode engine, "solve one" requests in an enum-mode engine; each step
// admits waiting requests into free slots (while slots and the device-memory
// budget allow), advances every in-flight request by one row and retires the
// ones that finished.  Requests that arrive while others are on the device
// join the same frontier at the next row.
struct d4_server {
    const d4_library *lib = nullptr;
    uint32_t slots = 0, max_h = 0;
    uint64_t budget = 0, est_per_req = 16ull << 20;
    std::unique_ptr<Engine> eng[2];  // [0] count, [1] enum
    std::vector<uint32_t> free_slots[2];
    std::vector<uint64_t> slot_req[2];
    struct Req {
        uint64_t id;
        std::vector<uint8_t> rows;
        bool single;
    };
    std::deque<Req> waiting;
    struct Done {
        uint64_t id, count;
        std::vector<uint8_t> sol;
    };
    std::deque<Done> done;
    uint64_t next_id = 1, admitted = 0, rejected_mem = 0;

    uint32_t in_flight() const { return eng[0]->n_active + eng[1]->n_active; }

    void admit() {
        while (!waiting.empty()) {
            Req &r = waiting.front();
            const int k = r.single ? 1 : 0;
            if (free_slots[k].empty()) break;
            // admission by device memory: what the engines hold now, plus
            // the larger of a default and the observed bytes per request
            const uint64_t used = eng[0]->device_bytes() + eng[1]->device_bytes();
            const uint32_t fl = in_flight();
            if (fl) est_per_req = std::max(est_per_req, used / fl);
            if (fl && used + est_per_req > budget) {
                rejected_mem++;
                break;
            }
            const uint32_t slot = free_slots[k].back();
            free_slots[k].pop_back();
            eng[k]->admit(slot, r.rows.data(), (uint32_t)r.rows.size());
            slot_req[k][slot] = r.id;
            admitted++;
            waiting.pop_front();
        }
    }

    // Every in-flight request advances one row.  The count engine (all
    // solutions) and the enum engine (first solution) step at the same time,
    // each from its own host thread on its own stream, so a step lasts as
    // long as the slower of the two rather than their sum.
    void step() {
        admit();
        bool stepped[2] = {false, false};
        std::exception_ptr err[2];
        std::map<uint32_t, std::vector<std::vector<uint8_t>>> firsts;
        auto advance = [&](int k) {
            try {
                HIPCHK(hipSetDevice(lib->device));
                Engine &e = *eng[k];
                e.step();
                if (k == 1 && !e.completed.empty()) firsts = e.first_solutions(e.completed);
            } catch (...) {
                err[k] = std::current_exception();
            }
        };
        for (int k = 0; k < 2; k++) stepped[k] = eng[k]->n_active != 0;
        if (stepped[0] && stepped[1]) {
            std::thread t(advance, 0);
            advance(1);
            t.join();
        } else {
            for (int k = 0; k < 2; k++)
                if (stepped[k]) advance(k);
        }
        // requests the other engine finished are delivered before an error
        // of one engine is raised
        for (int k = 0; k < 2; k++) {
            if (!stepped[k] || err[k]) continue;
            Engine &e = *eng[k];
            for (uint32_t slot : e.completed) {
                Done d{slot_req[k][slot], e.host_result[slot], {}};
                if (k == 1)
                    for (auto &sq : firsts[slot]) e.encode(slot, sq, d.sol);
                done.push_back(std::move(d));
                free_slots[k].push_back(slot);
            }
            e.trim_logs();
        }
        for (int k = 0; k < 2; k++)
            if (err[k]) std::rethrow_exception(err[k]);
    }
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

int d4_device_release_cached(uint64_t *bytes) {
    return guard([&] {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceSynchronize());
        const uint64_t n = pool().cached(dev);
        pool().trim(dev);
        if (bytes) *bytes = n;
    });
}

int d4_device_atomic_rate(double *ops_per_s) {
    return guard([&] {
        if (!ops_per_s) throw Error(D4_EINVAL, "null argument");
        const uint64_t m = 1ull << 27;  // 1 GiB of counters
        const uint32_t blocks = 8192, iters = 64;
        DevBuf<unsigned long long> tab;
        tab.reserve(m);
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 0;
        for (int rep = 0; rep < 3; rep++) {
            HIPCHK(hipEventRecord(e0, st));
            launch_atomic_rate(tab.p, m, iters, blocks, st);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms > 0) best = std::max(best, (double)blocks * 256 * iters / (ms * 1e-3));
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
        *ops_per_s = best;
    });
}

int d4_device_synchronize(void) {
    return guard([&] { HIPCHK(hipDeviceSynchronize()); });
}

size_t d4_free_polyominoes(unsigned n, uint64_t *out, size_t cap) {
    return d4_polyominoes(n, 0xffu, out, cap);
}

int d4_library_create(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                      uint32_t sym, uint32_t width, d4_library **out) {
    return d4_library_create_ex(shapes, counts, n, sym, nullptr, width, out);
}

int d4_library_create_ex(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                         uint32_t sym, const uint8_t *enabled, uint32_t width, d4_library **out) {
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
        build_tables(*L, shapes, counts, enabled);
        int dev = 0, ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (!ndev) throw Error(D4_EDEVICE, "no HIP device");
        HIPCHK(hipGetDevice(&dev));
        L->device = dev;
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        L->n_cu = prop.multiProcessorCount;
        L->lds_limit = (uint32_t)prop.sharedMemPerBlock;
        // single-library tables go to the device now (the allocation-free
        // entry point needs them there), and this device's scratch is set up
        if (L->ncand <= (uint32_t)MAX_CAND) ensure_fast(L.get());
        noalloc_setup(dev);
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

int d4_library_set_launch(d4_library *lib, uint32_t fit_rows, uint32_t threads,
                          uint32_t refill) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (refill < 1 || refill > 64) throw Error(D4_EINVAL, "refill must be 1..64");
        if (threads != 256 && threads != 512 && threads != 1024)
            throw Error(D4_EINVAL, "threads must be 256, 512 or 1024");
        if (fit_rows < 1 || fit_rows > lib->maxh) throw Error(D4_EINVAL, "fit_rows out of range");
        if (extend_lds_bytes(fit_rows, threads, false) > lib->lds_limit)
            throw Error(D4_EINVAL, "launch shape exceeds LDS");
        ensure_fast(lib);
        if (fit_rows != lib->fit_rows) {
            HIPCHK(hipSetDevice(lib->device));
            build_fit(*lib, fit_rows);
            uint64_t *old = lib->d_fit;
            lib->d_fit = nullptr;
            upload(lib->d_fit, lib->fit.data(), lib->fit.size());
            if (old) (void)hipFree(old);
        }
        lib->ext_threads = threads;
        lib->ext_refill = refill;
    });
}

int d4_library_set_budget(d4_library *lib, uint32_t steps) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        lib->ext_budget = steps;
    });
}

int d4_library_set_kernel(d4_library *lib, uint32_t kernel, uint32_t tail_rounds) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (kernel > D4_KERNEL_LEVEL) throw Error(D4_EINVAL, "unknown kernel");
        lib->xp_kernel = kernel;
        lib->xp_budget = tail_rounds;
    });
}

int d4_library_set_step_mode(d4_library *lib, uint32_t mode) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (mode > D4_STEP_GRAPH) throw Error(D4_EINVAL, "unknown step mode");
        lib->step_mode = mode;
    });
}

int d4_library_get_step_mode(const d4_library *lib, uint32_t *mode) {
    return guard([&] {
        if (!lib || !mode) throw Error(D4_EINVAL, "null argument");
        *mode = lib->step_mode;
    });
}

int d4_library_set_dedup(d4_library *lib, uint32_t mode) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (mode > D4_DEDUP_PART) throw Error(D4_EINVAL, "unknown dedup mode");
        lib->dedup_mode = mode;
    });
}

int d4_library_get_dedup(const d4_library *lib, uint32_t *mode) {
    return guard([&] {
        if (!lib || !mode) throw Error(D4_EINVAL, "null argument");
        *mode = lib->dedup_mode;
    });
}

int d4_library_set_donation_capacity(d4_library *lib, uint32_t items) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (items < 64) throw Error(D4_EINVAL, "donation capacity below 64 items");
        lib->don_items = items;
    });
}

int d4_library_get_kernel(const d4_library *lib, uint32_t *kernel, uint32_t *tail_rounds) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (kernel) *kernel = lib->xp_kernel;
        if (tail_rounds) *tail_rounds = lib->xp_budget;
    });
}

int d4_library_get_budget(const d4_library *lib, uint32_t *steps) {
    return guard([&] {
        if (!lib || !steps) throw Error(D4_EINVAL, "null argument");
        *steps = lib->ext_budget;
    });
}

int d4_library_get_launch(const d4_library *lib, uint32_t *fit_rows, uint32_t *threads,
                          uint32_t *refill) {
    return guard([&] {
        if (!lib) throw Error(D4_EINVAL, "null library");
        if (fit_rows) *fit_rows = lib->fit_rows;
        if (threads) *threads = lib->ext_threads;
        if (refill) *refill = lib->ext_refill;
    });
}

int d4_library_tune(d4_library *lib, uint32_t n_inst, const uint8_t *boards, uint32_t stride,
                    const uint32_t *heights, uint32_t *fit_rows, uint32_t *threads,
                    uint32_t *refill, double *best_ms) {
    return guard([&] {
        if (!lib || !boards || !heights || !n_inst) throw Error(D4_EINVAL, "null argument");
        uint32_t best_f = lib->fit_rows, best_t = lib->ext_threads, best_r = lib->ext_refill;
        double best = -1;
        auto measure = [&]() {
            auto e = make_batch(lib, n_inst, boards, stride, heights, D4_MODE_COUNT, 0);
            e->run();
            return e->stats.extend_ms;
        };
        measure();  // warm the allocator cache and the code objects
        // coordinate search: table rows x workgroup size first, then the
        // refill threshold at the best of those
        for (uint32_t f = 1; f <= std::min(lib->maxh, 4u); f++)
            for (uint32_t t : {256u, 512u, 1024u}) {
                if (extend_lds_bytes(f, t, false) > lib->lds_limit) continue;
                if (d4_library_set_launch(lib, f, t, best_r) != D4_OK) continue;
                double ms = measure();
                if (best < 0 || ms < best) {
                    best = ms;
                    best_f = f;
                    best_t = t;
                }
            }
        for (uint32_t r : {1u, 4u, 8u, 16u, 32u, 48u}) {
            if (d4_library_set_launch(lib, best_f, best_t, r) != D4_OK) continue;
            double ms = measure();
            if (ms < best) {
                best = ms;
                best_r = r;
            }
        }
        if (d4_library_set_launch(lib, best_f, best_t, best_r) != D4_OK)
            throw Error(D4_EDEVICE, g_err);
        // then the donation budget (work split for the launch's tail)
        uint32_t best_b = lib->ext_budget;
        for (uint32_t b : {0u, 512u, 768u, 1024u, 1536u, 2048u}) {
            lib->ext_budget = b;
            double ms = measure();
            if (ms < best) {
                best = ms;
                best_b = b;
            }
        }
        lib->ext_budget = best_b;
        if (fit_rows) *fit_rows = best_f;
        if (threads) *threads = best_t;
        if (refill) *refill = best_r;
        if (best_ms) *best_ms = best;
    });
}

int d4_searcher_create(const d4_library *lib, const uint8_t *rows, uint32_t height,
                       uint32_t mode, uint32_t flags, d4_searcher **out) {
    return guard([&] {
        if (!lib || !out || (!rows && height)) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        auto s = std::make_unique<d4_searcher>();
        s->e = make_batch(lib, 1, rows, height, &height, mode, flags);
        if (!height) s->e->run();  // an empty board has its one (empty) tiling
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

int d4_searcher_set_memory(d4_searcher *s, uint64_t bytes) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        if (s->e->mode != D4_MODE_COUNT) throw Error(D4_ESTATE, "memory cap needs count mode");
        if (bytes < (1u << 20)) throw Error(D4_EINVAL, "memory cap below 1 MiB");
        s->e->mem_cap = bytes;
    });
}

int d4_searcher_set_host_memory(d4_searcher *s, uint64_t bytes) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        if (s->e->mode != D4_MODE_COUNT) throw Error(D4_ESTATE, "host parking needs count mode");
        if (s->e->parked.size()) throw Error(D4_ESTATE, "parts are parked");
        s->e->host_cap = bytes;
    });
}

void d4_searcher_destroy(d4_searcher *s) { delete s; }

int d4_searcher_count_noalloc(const d4_library *lib, const uint8_t *rows, uint32_t height,
                              uint32_t flags, uint64_t *count) {
    return guard([&] {
        if (!lib || !count || (!rows && height)) throw Error(D4_EINVAL, "null argument");
        *count = 0;
        if (height > (uint32_t)MAX_ROWS) throw Error(D4_ERANGE, "board taller than 64 rows");
        if (!height) {
            *count = 1;  // the empty board has its one (empty) tiling
            return;
        }
        const int dev = lib->device;
        if (dev < 0 || dev >= kMaxDevices || !g_noalloc[dev].ready)
            throw Error(D4_ESTATE, "no allocation-free scratch on the library's device");
        NoAllocDev &na = g_noalloc[dev];
        std::lock_guard<std::mutex> lk(na.mu);
        HIPCHK(hipSetDevice(dev));
        Engine e(lib, height, flags, na);
        e.admit(0, rows, height);
        e.run();
        *count = e.host_result[0];
    });
}

uint32_t d4_searcher_height(const d4_searcher *s) { return s ? s->e->height_of(0) : 0; }

uint64_t d4_searcher_size(const d4_searcher *s) { return s ? s->e->n_front : 0; }

uint64_t d4_searcher_next_size(const d4_searcher *s) {
    uint64_t v = 0;
    if (!s) return 0;
    if (guard([&] { v = s->e->next_size(); }) != D4_OK) return 0;
    return v;
}

int d4_searcher_set_sample(d4_searcher *s, uint32_t height, uint64_t cap) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        s->e->sample_h = height;
        s->e->sample_cap = cap;
        s->e->sample_keys.clear();
        s->e->sample_cnt.clear();
        s->e->take_sample();
    });
}

int d4_searcher_sample(const d4_searcher *s, uint64_t *keys, uint64_t *counts, uint64_t cap,
                       uint64_t *n) {
    return guard([&] {
        if (!s || !n) throw Error(D4_EINVAL, "null argument");
        const Engine &e = *s->e;
        *n = e.sample_keys.size();
        for (uint64_t k = 0; k < *n && k < cap; k++) {
            if (keys) keys[k] = e.sample_keys[k];
            if (counts) counts[k] = e.sample_cnt[k];
        }
    });
}

int d4_searcher_frontier(const d4_searcher *s, uint64_t first, uint64_t cap, uint64_t *keys,
                         uint64_t *counts, uint64_t *n) {
    return guard([&] {
        if (!s || !n) throw Error(D4_EINVAL, "null argument");
        Engine &e = *s->e;
        e.upload_pending();
        *n = e.n_front;
        if (first >= e.n_front || !cap) return;
        const uint64_t k = std::min<uint64_t>(cap, e.n_front - first);
        if (keys) e.copy(keys, e.fk.p + first, 8 * k, hipMemcpyDeviceToHost);
        if (counts) e.copy(counts, e.fc.p + first, 8 * k, hipMemcpyDeviceToHost);
    });
}

uint64_t d4_searcher_size_bucket(const d4_searcher *s, unsigned ea) {
    uint64_t v = 0;
    if (!s || ea > 255) return 0;
    if (guard([&] {
            s->e->buckets();
            v = s->e->bucket_n[ea];
        }) != D4_OK)
        return 0;
    return v;
}

uint64_t d4_searcher_next_size_bucket(const d4_searcher *s, unsigned ea) {
    uint64_t v = 0;
    if (!s || ea > 255) return 0;
    if (guard([&] {
            s->e->buckets();
            v = s->e->bucket_next[ea];
        }) != D4_OK)
        return 0;
    return v;
}

int d4_searcher_step(d4_searcher *s) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null searcher");
        if (!s->e->n_active) throw Error(D4_ESTATE, "search already finished");
        s->e->step();
        // after the last row the terminal states are absorbed in the same call
        if (s->e->n_active && s->e->height_of(0) == 0) s->e->step();
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
        if (s->e->n_active) throw Error(D4_ESTATE, "search not finished");
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

int d4_searcher_export(const d4_searcher *s, void *rx, size_t cap, size_t *n,
                       uint64_t *counts) {
    return guard([&] {
        if (!s || !n) throw Error(D4_EINVAL, "null argument");
        Engine &e = *s->e;
        const d4_library &L = *e.lib;
        if (!e.active[0]) throw Error(D4_ESTATE, "search finished: no frontier left");
        e.upload_pending();
        const uint32_t H = e.height_of(0);
        if (H > 8 || H == 0) throw Error(D4_ERANGE, "the record format holds 1..8 rows");
        if (L.n > 16) throw Error(D4_ERANGE, "the record format holds at most 16 pieces");
        for (uint32_t i = 0; i < L.n; i++)
            if (L.counts[i] > 1) throw Error(D4_ERANGE, "the record format has no piece counts");
        std::vector<uint64_t> keys(e.n_front), cnts(e.n_front);
        if (e.n_front) {
            e.copy(keys.data(), e.fk.p, e.n_front * 8, hipMemcpyDeviceToHost);
            e.copy(cnts.data(), e.fc.p, e.n_front * 8, hipMemcpyDeviceToHost);
        }
        const uint32_t row = e.g - e.start[0];
        uint64_t BW = 0;
        for (uint32_t j = 0; j < 8; j++)
            BW |= (uint64_t)e.boards[row + j] << (8 * j);
        struct Rec {
            uint32_t w[6];  // R (5 words) + ea in the low byte of w[5]
            uint64_t cnt;
        };
        std::vector<Rec> out(e.n_front);
        for (size_t i = 0; i < keys.size(); i++) {
            const uint64_t E = BW & ~(keys[i] & e.occ_mask);
            const uint32_t used = (uint32_t)(keys[i] >> e.occ_bits) & e.used_mask;
            uint8_t nm[16];
            std::memset(nm, 0xff, sizeof nm);
            uint32_t cnt = 0;
            for (uint32_t p = 0; p < L.n; p++)  // ascending ids = sorted list
                if ((used >> L.off[p]) & 1u) nm[cnt++] = (uint8_t)p;
            const uint32_t cap_ids = H == 8 ? 0 : H == 7 ? 8 : H == 6 ? 14 : H == 5 ? 15 : 16;
            if (cnt > cap_ids) throw Error(D4_ERANGE, "more pieces than the record holds");
            uint32_t ex[4];
            for (int w = 0; w < 4; w++)
                ex[w] = nm[4 * w] | nm[4 * w + 1] << 8 | nm[4 * w + 2] << 16 |
                        (uint32_t)nm[4 * w + 3] << 24;
            const uint64_t area = E >> 8;  // rows 1.. of the window
            Rec &r = out[i];
            r.w[2] = ex[0];
            r.w[3] = ex[1];
            r.w[4] = ex[2];
            if (H == 8) {  // the root record: CudaSearcher's ctor leaves ex0..2 zero
                r.w[0] = (uint32_t)area;
                r.w[1] = (uint32_t)(area >> 32);
                r.w[2] = r.w[3] = r.w[4] = 0;
            } else if (H == 7) {  // ex2 is not part of a height-7 record
                r.w[0] = (uint32_t)area;
                r.w[1] = (uint32_t)(area >> 32 & 0xffffu) | cnt << 24;
                r.w[4] = ~0u;
            } else if (H == 6) {
                r.w[0] = (uint32_t)area;
                r.w[1] = (ex[3] & 0xffffu) | (uint32_t)(area >> 32 & 0xffu) << 16 | cnt << 24;
            } else if (H == 5) {
                r.w[0] = (uint32_t)area;
                r.w[1] = (ex[3] & 0xffffffu) | cnt << 24;
            } else {
                r.w[0] = (uint32_t)(area & 0xffffffu) | cnt << 24;
                r.w[1] = ex[3];
            }
            r.w[5] = (uint32_t)(E & 0xffu);
            r.cnt = cnts[i];
        }
        std::sort(out.begin(), out.end(), [](const Rec &a, const Rec &b) {
            if (a.w[5] != b.w[5]) return a.w[5] < b.w[5];
            return std::lexicographical_compare(a.w, a.w + 5, b.w, b.w + 5);
        });
        *n = out.size();
        if (!rx || !cap) return;
        if (cap < out.size()) throw Error(D4_ERANGE, "record buffer too small");
        uint8_t *dst = static_cast<uint8_t *>(rx);
        for (size_t i = 0; i < out.size(); i++) {
            std::memcpy(dst + 24 * i, out[i].w, 24);
            if (counts) counts[i] = out[i].cnt;
        }
    });
}

int d4_searcher_stats(const d4_searcher *s, d4_stats *st) {
    return guard([&] {
        if (!s || !st) throw Error(D4_EINVAL, "null argument");
        *st = s->e->stats;
    });
}

int d4_server_create(const d4_library *lib, uint32_t slots, uint32_t max_height,
                     uint64_t mem_budget, d4_server **out) {
    return guard([&] {
        if (!lib || !out) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        if (!slots) throw Error(D4_EINVAL, "need at least one slot");
        auto sv = std::make_unique<d4_server>();
        sv->lib = lib;
        sv->slots = slots;
        sv->max_h = max_height;
        HIPCHK(hipSetDevice(lib->device));
        if (!mem_budget) mem_budget = available_device_bytes() / 2;
        sv->budget = mem_budget;
        for (int k = 0; k < 2; k++) {
            sv->eng[k] = std::make_unique<Engine>(lib, slots, max_height,
                                                  k ? D4_MODE_ENUM : D4_MODE_COUNT, 0, false);
            sv->slot_req[k].assign(slots, 0);
            for (uint32_t i = slots; i-- > 0;) sv->free_slots[k].push_back(i);
        }
        *out = sv.release();
    });
}

void d4_server_destroy(d4_server *s) { delete s; }

int d4_server_submit(d4_server *s, const uint8_t *rows, uint32_t height, int single,
                     uint64_t *req_id) {
    return guard([&] {
        if (!s || (!rows && height)) throw Error(D4_EINVAL, "null argument");
        if (height > s->max_h) throw Error(D4_EINVAL, "board taller than the server's maximum");
        for (uint32_t r = 0; r < height; r++)
            if (s->lib->W < 8 && (rows[r] >> s->lib->W))
                throw Error(D4_EINVAL, "board wider than the library's width");
        d4_server::Req r{s->next_id++, std::vector<uint8_t>(rows, rows + height), single != 0};
        if (req_id) *req_id = r.id;
        s->waiting.push_back(std::move(r));
    });
}

int d4_server_step(d4_server *s, uint32_t *in_flight, uint32_t *waiting) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null server");
        s->step();
        if (in_flight) *in_flight = s->in_flight();
        if (waiting) *waiting = (uint32_t)s->waiting.size();
    });
}

int d4_server_poll(d4_server *s, uint64_t *req_id, uint64_t *count, uint8_t *sol, size_t cap,
                   size_t *len) {
    int got = 0;
    int rc = guard([&] {
        if (!s) throw Error(D4_EINVAL, "null server");
        if (s->done.empty()) return;
        d4_server::Done &d = s->done.front();
        if (len) *len = d.sol.size();
        if (sol && d.sol.size() > cap) throw Error(D4_ERANGE, "solution buffer too small");
        if (sol && !d.sol.empty()) std::memcpy(sol, d.sol.data(), d.sol.size());
        if (req_id) *req_id = d.id;
        if (count) *count = d.count;
        s->done.pop_front();
        got = 1;
    });
    return rc == D4_OK ? got : rc;
}

int d4_server_stats(const d4_server *s, d4_stats *all, d4_stats *one, uint64_t *admitted,
                    uint64_t *mem_waits, uint64_t *device_bytes) {
    return guard([&] {
        if (!s) throw Error(D4_EINVAL, "null server");
        if (all) *all = s->eng[0]->stats;
        if (one) *one = s->eng[1]->stats;
        if (admitted) *admitted = s->admitted;
        if (mem_waits) *mem_waits = s->rejected_mem;
        if (device_bytes) *device_bytes = s->eng[0]->device_bytes() + s->eng[1]->device_bytes();
    });
}

int d4_batch_create(const d4_library *lib, uint32_t n_inst, const uint8_t *boards,
                    uint32_t stride, const uint32_t *heights, uint32_t mode, uint32_t flags,
                    d4_batch **out) {
    return guard([&] {
        if (!lib || !out || !boards || !heights) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        auto b = std::make_unique<d4_batch>();
        b->e = make_batch(lib, n_inst, boards, stride, heights, mode, flags);
        *out = b.release();
    });
}

void d4_batch_destroy(d4_batch *b) { delete b; }

int d4_batch_set_memory(d4_batch *b, uint64_t bytes) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (b->e->mode != D4_MODE_COUNT) throw Error(D4_ESTATE, "memory cap needs count mode");
        if (bytes < (1u << 20)) throw Error(D4_EINVAL, "memory cap below 1 MiB");
        b->e->mem_cap = bytes;
    });
}

int d4_batch_set_host_memory(d4_batch *b, uint64_t bytes) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        if (b->e->mode != D4_MODE_COUNT) throw Error(D4_ESTATE, "host parking needs count mode");
        if (b->e->parked.size()) throw Error(D4_ESTATE, "parts are parked");
        b->e->host_cap = bytes;
    });
}

int d4_batch_run(d4_batch *b) {
    return guard([&] {
        if (!b) throw Error(D4_EINVAL, "null batch");
        b->e->run();
    });
}

int d4_batch_counts(const d4_batch *b, uint64_t *counts) {
    return guard([&] {
        if (!b || !counts) throw Error(D4_EINVAL, "null argument");
        if (b->e->n_active) throw Error(D4_ESTATE, "batch not finished");
        for (uint32_t i = 0; i < b->e->S; i++) counts[i] = b->e->host_result[i];
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
