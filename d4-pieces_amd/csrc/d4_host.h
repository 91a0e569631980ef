// d4_host.h -- host-side helpers shared by the engines: error plumbing,
// bitboard geometry and the caching device allocator.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "d4_internal.h"

namespace d4 {

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
inline int guard(F &&f) {
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

// A row step's unique states before compaction: the compaction writes them
// at 32-bit offsets, so a row with 2^32 or more is refused before it runs.
inline void check_row_states(uint64_t unique) {
    if (unique > 0xffffffffull)
        throw Error(D4_ERANGE, "a row's unique states (" + std::to_string(unique) +
                                   ") exceed the compaction's 2^32");
}

// The number of states a compaction found (c32[C32_OUT]) must equal the
// step's new-slot count, compared in 64 bits; otherwise the table and the
// counts disagree.
inline void check_compacted(uint64_t found, uint64_t expected) {
    if (found != expected)
        throw Error(D4_EDEVICE, "internal error: compaction found " + std::to_string(found) +
                                    " states, the row step counted " +
                                    std::to_string(expected));
}

// --------------------------------------------------------------------------
// geometry (8-wide bitboards, bit = row*8 + col)
// --------------------------------------------------------------------------

inline uint64_t normalize8(uint64_t v) {
    if (!v) return 0;
    while (!(v & 0xffull)) v >>= 8;
    while (!(v & 0x0101010101010101ull)) v >>= 1;
    return v;
}

// D4 transform t in the reference's order (Shape.cpp:28-40: identity, flip
// X, flip Y, both, then the same three after a transpose).  Written as the
// image of each input cell: flips act on the input's coordinates, then the
// transpose swaps them.
inline uint64_t transform8(uint64_t s, unsigned t) {
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

inline unsigned width8(uint64_t s) {
    unsigned w = 0;
    for (unsigned c = 0; c < 8; c++)
        if (s & (0x0101010101010101ull << c)) w = c + 1;
    return w;
}

inline unsigned height8(uint64_t s) {
    unsigned h = 0;
    for (unsigned r = 0; r < 8; r++)
        if (s & (0xffull << (8 * r))) h = r + 1;
    return h;
}

inline unsigned bitw(uint32_t v) { return v ? 32u - (unsigned)__builtin_clz(v) : 0u; }

inline uint64_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Slots of a single-library dedup table: at least 1024, a multiple of 1024
// (the kernels map hashes onto any slot count).
inline uint64_t round_slots(uint64_t v) {
    return std::max<uint64_t>((v + 1023) & ~1023ull, 1024);
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
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = hipMalloc(&p, bytes);
        bool trimmed = false;
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(mu_);
            trim_locked();
            trimmed = true;
            e = hipMalloc(&p, bytes);
        }
        static const bool trace = getenv("D4_TRACE") != nullptr;
        if (trace)
            fprintf(stderr, "pool: hipMalloc %.3f GiB in %.1f ms%s\n", (double)bytes / (1u << 30),
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                        .count(),
                    trimmed ? " (cache released first)" : "");
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
    // Return the blocks cached on a device to the runtime (its pending work
    // must be finished).
    void trim(int dev) {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = free_.begin(); it != free_.end();) {
            if (it->second.first == dev) {
                int cur = 0;
                (void)hipGetDevice(&cur);
                (void)hipSetDevice(dev);
                (void)hipFree(it->second.second);
                (void)hipSetDevice(cur);
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
    }
    // Bytes cached for reuse on a device: free for the library's purposes,
    // though hipMemGetInfo counts them as in use.
    size_t cached(int dev) {
        std::lock_guard<std::mutex> lk(mu_);
        size_t n = 0;
        for (auto &kv : free_)
            if (kv.second.first == dev) n += kv.first;
        return n;
    }
};

inline Pool &pool() {
    static Pool *p = new Pool();  // never destroyed: outlives every handle
    return *p;
}

// Device memory a new search may plan with: hipMemGetInfo's free bytes plus
// what the allocator holds cached on the current device.
inline uint64_t available_device_bytes() {
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    return (uint64_t)free_b + (uint64_t)pool().cached(dev);
}

// A device buffer from the caching allocator, or attached to a fixed slice
// of memory the caller owns (the allocation-free entry point's scratch):
// an attached buffer never allocates or frees, and a request beyond its
// slice is an error.
template <class T>
struct DevBuf {
    T *p = nullptr;
    uint64_t cap = 0;
    bool ext = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void attach(T *q, uint64_t n) {
        release();
        p = q;
        cap = n;
        ext = true;
    }
    void release() {
        if (ext) return;
        if (p) pool().release(p, cap * sizeof(T));
        p = nullptr;
        cap = 0;
    }
    void reserve(uint64_t n) {
        if (n <= cap) return;
        if (ext)
            throw Error(D4_ERANGE, "allocation-free search: " + std::to_string(n) +
                                       " entries needed, the scratch slice holds " +
                                       std::to_string(cap));
        release();
        uint64_t c = Pool::round(std::max<uint64_t>(n, 64) * sizeof(T)) / sizeof(T);
        p = static_cast<T *>(pool().alloc(c * sizeof(T)));
        cap = c;
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(ext, o.ext);
    }
};

}  // namespace d4
