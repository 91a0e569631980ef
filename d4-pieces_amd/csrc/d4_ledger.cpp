// d4_ledger.cpp -- the consumer of terminal states: piece-set names and
// the ledgers they are logged into.  Host code only (no device calls): a
// batch search hands its terminal states over as (instance, counters)
// pairs, they are named here and logged per board configuration.
//
// Names follow src/naming.hpp:15-71 / naming.inl:6-72 / naming.cpp:9-81;
// ledgers follow src/search.cpp:25-226.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <memory>
#include <unordered_set>
#include <vector>

#include "d4_host.h"
#include "d4_library.h"

using namespace d4;

namespace {

// a + b and a * b that refuse to wrap (the names must stay unique)
uint64_t add_chk(uint64_t a, uint64_t b) {
    uint64_t r;
    if (__builtin_add_overflow(a, b, &r)) throw Error(D4_ERANGE, "naming: count exceeds 2^64");
    return r;
}
uint64_t mul_chk(uint64_t a, uint64_t b) {
    uint64_t r;
    if (__builtin_mul_overflow(a, b, &r)) throw Error(D4_ERANGE, "naming: count exceeds 2^64");
    return r;
}

}  // namespace

struct d4_naming {
    uint64_t min_m = 0, max_m = 0, min_n = 0, max_n = 0, tgt = 0;
    uint32_t D = 0;                   // size classes: m = min_m + d, d < D
    std::vector<uint64_t> cls;        // pieces per size class
    std::vector<uint64_t> first;      // id of the class's first piece; first[D] = total
    std::vector<std::vector<uint64_t>> choose;  // choose[s][k] = C(s, k), s <= max class
    // partitions: pieces taken per class, ascending in lexicographic order
    std::vector<std::vector<uint32_t>> parts;
    // per partition: C(class size, k) per class and the product of the
    // entries after each class (mixed radix, class 0 most significant)
    std::vector<std::vector<uint64_t>> digits, below;
    std::vector<uint64_t> start;  // first name of each partition; start.back() = size

    uint64_t binom(uint64_t s, uint64_t k) const { return k > s ? 0 : choose[s][k]; }

    void build() {
        D = (uint32_t)(max_m - min_m + 1);
        first.assign(D + 1, 0);
        for (uint32_t d = 0; d < D; d++) first[d + 1] = add_chk(first[d], cls[d]);
        const uint64_t smax = cls.empty() ? 0 : *std::max_element(cls.begin(), cls.end());
        if (smax > (1u << 16)) throw Error(D4_EINVAL, "naming: more than 65536 pieces of one size");
        // Pascal's triangle; entries past 2^64 are marked and refused when used
        choose.assign(smax + 1, {});
        for (uint64_t s = 0; s <= smax; s++) {
            choose[s].assign(s + 1, 1);
            for (uint64_t k = 1; k < s; k++) {
                uint64_t v;
                choose[s][k] = __builtin_add_overflow(choose[s - 1][k - 1], choose[s - 1][k], &v)
                                   ? UINT64_MAX
                                   : v;
            }
        }
        // Every vector k (0 <= k[d] <= cls[d]) with sum k[d] * (min_m + d) ==
        // tgt and min_n <= sum k[d] <= max_n, produced in lexicographic order
        // (smallest k[0] first, and so on down the classes).
        std::vector<uint32_t> k(D, 0);
        struct Rec {
            d4_naming &N;
            std::vector<uint32_t> &k;
            void go(uint32_t d, uint64_t left, uint64_t n) {
                if (!left) {
                    if (n >= N.min_n) N.parts.push_back(k);
                    return;
                }
                if (d == N.D) return;
                const uint64_t m = N.min_m + d;
                for (uint64_t c = 0; c * m <= left && n + c <= N.max_n && c <= N.cls[d]; c++) {
                    k[d] = (uint32_t)c;
                    go(d + 1, left - c * m, n + c);
                }
                k[d] = 0;
            }
        } rec{*this, k};
        rec.go(0, tgt, 0);
        start.assign(1, 0);
        for (const auto &p : parts) {
            std::vector<uint64_t> dig(D), bel(D);
            uint64_t prod = 1;
            for (uint32_t d = D; d-- > 0;) {
                bel[d] = prod;
                dig[d] = binom(cls[d], p[d]);
                if (dig[d] == UINT64_MAX) throw Error(D4_ERANGE, "naming: count exceeds 2^64");
                prod = mul_chk(prod, dig[d]);
            }
            digits.push_back(std::move(dig));
            below.push_back(std::move(bel));
            start.push_back(add_chk(start.back(), prod));
        }
    }

    uint64_t size() const { return start.back(); }
    uint64_t size_pieces() const { return first[D]; }

    // Rank of the chosen subset of class d (k of cls[d] pieces) among the
    // k-subsets in lexicographic order of their sorted indices: before a
    // subset that skips index i come all those that take it.
    template <class Used>
    uint64_t rank_subset(uint32_t d, uint64_t k, Used &&used) const {
        uint64_t r = 0, s = cls[d];
        for (uint64_t i = 0; k; i++, s--) {
            if (used(d, i))
                k--;
            else
                r += binom(s - 1, k - 1);
        }
        return r;
    }

    template <class Take>
    void unrank_subset(uint32_t d, uint64_t k, uint64_t r, Take &&take) const {
        uint64_t s = cls[d];
        for (uint64_t i = 0; k; i++, s--) {
            const uint64_t with_i = binom(s - 1, k - 1);
            if (r < with_i) {
                take(d, i);
                k--;
            } else {
                r -= with_i;
            }
        }
    }

    template <class Used>
    uint64_t name(Used &&used) const {
        std::vector<uint32_t> k(D, 0);
        for (uint32_t d = 0; d < D; d++)
            for (uint64_t i = 0; i < cls[d]; i++) k[d] += used(d, i) ? 1u : 0u;
        auto it = std::lower_bound(parts.begin(), parts.end(), k);
        if (it == parts.end() || *it != k) return D4_NO_NAME;
        const size_t p = (size_t)(it - parts.begin());
        uint64_t r = 0;
        for (uint32_t d = 0; d < D; d++) r = r * digits[p][d] + rank_subset(d, k[d], used);
        return start[p] + r;
    }

    template <class Take>
    bool resolve(uint64_t nm, Take &&take) const {
        if (nm >= size()) return false;
        const size_t p = (size_t)(std::upper_bound(start.begin(), start.end(), nm) - start.begin()) - 1;
        uint64_t r = nm - start[p];
        for (uint32_t d = 0; d < D; d++) {
            unrank_subset(d, parts[p][d], r / below[p][d], take);
            r %= below[p][d];
        }
        return true;
    }
};

// --------------------------------------------------------------------------
// ledgers
// --------------------------------------------------------------------------

struct d4_ledger {
    int kind = D4_LEDGER_VERIFY;
    uint64_t n_configs = 0, n_names = 0, row_bytes = 0;
    int fd = -1;
    void *map = nullptr;
    size_t map_bytes = 0;
    // file kinds: views into the mapping
    struct Desc {
        uint64_t shape, count, solved, padding;
    };
    static_assert(sizeof(Desc) == 32, "descriptor layout");
    Desc *cfgs = nullptr;
    uint64_t *nmes = nullptr;
    uint8_t *ledger = nullptr;
    // AtomicCount: per-name totals and per-config names already logged
    std::vector<uint64_t> counts;
    std::vector<std::unordered_set<uint64_t>> seen;
    uint64_t work = 0, configs_done = 0;

    ~d4_ledger() {
        if (map) (void)munmap(map, map_bytes);
        if (fd >= 0) (void)close(fd);
    }

    void open_file(const char *path) {
        row_bytes = kind == D4_LEDGER_BYTE ? n_configs : (n_configs + 7) / 8;
        const uint64_t head = mul_chk(n_configs, 32), names = mul_chk(n_names, 8);
        const uint64_t total = add_chk(add_chk(head, names), mul_chk(n_names, row_bytes));
        fd = ::open(path, O_RDWR | O_CREAT, 0666);
        if (fd < 0) throw Error(D4_EINVAL, std::string("open: ") + strerror(errno));
        struct stat sb;
        if (fstat(fd, &sb)) throw Error(D4_EINVAL, std::string("fstat: ") + strerror(errno));
        if ((uint64_t)sb.st_size > total) throw Error(D4_EINVAL, "ledger file too long");
        if (sb.st_size && (uint64_t)sb.st_size < total) throw Error(D4_EINVAL, "ledger file too short");
        if (!sb.st_size && ftruncate(fd, (off_t)total))
            throw Error(D4_EINVAL, std::string("ftruncate: ") + strerror(errno));
        map_bytes = total;
        if (total) {
            map = mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
            if (map == MAP_FAILED) {
                map = nullptr;
                throw Error(D4_ENOMEM, std::string("mmap: ") + strerror(errno));
            }
        }
        auto *b = static_cast<uint8_t *>(map);
        cfgs = reinterpret_cast<Desc *>(b);
        nmes = reinterpret_cast<uint64_t *>(b + head);
        ledger = b + head + names;
    }

    // one (config, name) pair; true when new for the config
    bool log(uint64_t c, uint64_t v) {
        if (v == D4_NO_NAME) return false;
        if (c >= n_configs || v >= n_names) throw Error(D4_EINVAL, "config or name out of range");
        switch (kind) {
            case D4_LEDGER_VERIFY:
                break;
            case D4_LEDGER_COUNT:
                if (!seen[c].insert(v).second) return false;
                counts[v]++;
                break;
            case D4_LEDGER_BYTE: {
                uint8_t &l = ledger[v * row_bytes + c];
                if (l) return false;
                l = 0xff;
                nmes[v]++;
                break;
            }
            default: {
                uint8_t &l = ledger[v * row_bytes + c / 8];
                const uint8_t bit = (uint8_t)(1u << (c % 8));
                if (l & bit) return false;
                l |= bit;
                nmes[v]++;
                break;
            }
        }
        work++;
        return true;
    }
};

extern "C" {

size_t d4_polyominoes(unsigned n, uint32_t sym, uint64_t *out, size_t cap) {
    if (n < 1 || n > 8 || !(sym & 0xffu)) return 0;
    // canonical form: the minimum over the group's transforms
    auto canon = [&](uint64_t s) {
        uint64_t v = ~0ull;
        for (unsigned t = 0; t < 8; t++)
            if (sym >> t & 1u) v = std::min(v, transform8(s, t));
        return v;
    };
    std::vector<uint64_t> cur{1};
    for (unsigned k = 2; k <= n; k++) {
        // every k-cell polyomino is a (k-1)-cell one plus a cell; growing
        // the canonical form covers the whole orbit, as the group maps the
        // grown shape onto a growth of the canonical form
        std::vector<uint64_t> nxt;
        for (uint64_t s0 : cur) {
            const uint64_t s = s0 << 9;  // a free row above, a free column left
            for (uint64_t b = s; b; b &= b - 1) {
                const int i = __builtin_ctzll(b), r = i / 8, c = i % 8;
                const int nb[4][2] = {{r - 1, c}, {r + 1, c}, {r, c - 1}, {r, c + 1}};
                for (auto &q : nb) {
                    if (q[0] < 0 || q[0] > 7 || q[1] < 0 || q[1] > 7) continue;
                    const uint64_t bit = 1ull << (q[0] * 8 + q[1]);
                    if (!(s & bit)) nxt.push_back(canon(s | bit));
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

int d4_naming_create(uint64_t min_m, uint64_t max_m, uint64_t min_n, uint64_t max_n,
                     uint64_t tgt, const uint64_t *arr_sizes, d4_naming **out) {
    return guard([&] {
        if (!out || !arr_sizes) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        if (min_m < 1 || min_m > max_m || max_m > 64)
            throw Error(D4_EINVAL, "naming: need 1 <= min_m <= max_m <= 64");
        auto N = std::make_unique<d4_naming>();
        N->min_m = min_m;
        N->max_m = max_m;
        N->min_n = min_n;
        N->max_n = max_n;
        N->tgt = tgt;
        for (uint64_t m = min_m; m <= max_m; m++) N->cls.push_back(arr_sizes[m]);
        N->build();
        *out = N.release();
    });
}

void d4_naming_destroy(d4_naming *nm) { delete nm; }

uint64_t d4_naming_size(const d4_naming *nm) { return nm ? nm->size() : 0; }

uint64_t d4_naming_size_pieces(const d4_naming *nm) { return nm ? nm->size_pieces() : 0; }

uint64_t d4_naming_name_piece(const d4_naming *nm, uint64_t m, uint64_t i) {
    if (!nm || m < nm->min_m || m > nm->max_m) return D4_NO_NAME;
    const uint32_t d = (uint32_t)(m - nm->min_m);
    return i < nm->cls[d] ? nm->first[d] + i : D4_NO_NAME;
}

int d4_naming_resolve_piece(const d4_naming *nm, uint64_t id, uint64_t *m, uint64_t *i) {
    return guard([&] {
        if (!nm || !m || !i) throw Error(D4_EINVAL, "null argument");
        *m = *i = 0;
        if (id >= nm->size_pieces()) throw Error(D4_ERANGE, "piece id out of range");
        const uint32_t d =
            (uint32_t)(std::upper_bound(nm->first.begin(), nm->first.end(), id) - nm->first.begin()) - 1;
        *m = nm->min_m + d;
        *i = id - nm->first[d];
    });
}

int d4_naming_name(const d4_naming *nm, const uint8_t *used, uint64_t *name) {
    return guard([&] {
        if (!nm || !used || !name) throw Error(D4_EINVAL, "null argument");
        *name = nm->name([&](uint32_t d, uint64_t i) { return used[nm->first[d] + i] != 0; });
    });
}

int d4_naming_resolve(const d4_naming *nm, uint64_t name, uint8_t *used) {
    return guard([&] {
        if (!nm || !used) throw Error(D4_EINVAL, "null argument");
        std::fill(used, used + nm->size_pieces(), 0);
        if (!nm->resolve(name, [&](uint32_t d, uint64_t i) { used[nm->first[d] + i] = 1; }))
            throw Error(D4_ERANGE, "name out of range");
    });
}

int d4_naming_name_counters(const d4_naming *nm, const d4_library *lib, const uint64_t *counters,
                            size_t n, uint64_t *names) {
    return guard([&] {
        if (!nm || !lib || (n && (!counters || !names))) throw Error(D4_EINVAL, "null argument");
        const uint64_t P = nm->size_pieces();
        if (lib->n != P) throw Error(D4_EINVAL, "library pieces differ from the naming's");
        // piece id -> (class, index): ids are laid out class by class
        std::vector<uint8_t> used(P);
        for (size_t k = 0; k < n; k++) {
            bool once = true;
            for (uint32_t p = 0; p < (uint32_t)P; p++) {
                const uint64_t c = (counters[k] >> lib->off[p]) & ((1ull << lib->fbits[p]) - 1ull);
                used[p] = c != 0;
                once &= c <= 1;
            }
            names[k] = once ? nm->name([&](uint32_t d, uint64_t i) { return used[nm->first[d] + i] != 0; })
                            : D4_NO_NAME;
        }
    });
}

int d4_library_counters(const d4_library *lib, uint32_t *offset, uint32_t *width, uint32_t *n) {
    return guard([&] {
        if (!lib || !n) throw Error(D4_EINVAL, "null argument");
        *n = lib->n;
        for (uint32_t p = 0; p < lib->n; p++) {
            if (offset) offset[p] = lib->off[p];
            if (width) width[p] = lib->fbits[p];
        }
    });
}

int d4_ledger_open(int kind, const char *path, uint64_t n_configs, uint64_t n_names,
                   d4_ledger **out) {
    return guard([&] {
        if (!out) throw Error(D4_EINVAL, "null argument");
        *out = nullptr;
        if (kind < D4_LEDGER_VERIFY || kind > D4_LEDGER_BYTE) throw Error(D4_EINVAL, "unknown ledger kind");
        auto L = std::make_unique<d4_ledger>();
        L->kind = kind;
        L->n_configs = n_configs;
        L->n_names = n_names;
        if (kind == D4_LEDGER_COUNT) {
            L->counts.assign(n_names, 0);
            L->seen.resize(n_configs);
        } else if (kind != D4_LEDGER_VERIFY) {
            if (!path) throw Error(D4_EINVAL, "file ledgers need a path");
            L->open_file(path);
        }
        *out = L.release();
    });
}

void d4_ledger_close(d4_ledger *L) { delete L; }

int d4_ledger_should_run(d4_ledger *L, uint64_t config, uint64_t shape, int *run) {
    return guard([&] {
        if (!L || !run) throw Error(D4_EINVAL, "null argument");
        if (config >= L->n_configs) throw Error(D4_EINVAL, "config out of range");
        *run = 1;
        if (L->cfgs) {
            L->cfgs[config].shape = shape;
            *run = L->cfgs[config].solved == 0;
        }
    });
}

int d4_ledger_log(d4_ledger *L, uint64_t config, uint64_t name, int *fresh) {
    return guard([&] {
        if (!L) throw Error(D4_EINVAL, "null argument");
        const bool f = L->log(config, name);
        if (fresh) *fresh = f;
    });
}

int d4_ledger_log_many(d4_ledger *L, const uint64_t *configs, const uint64_t *names, size_t n,
                       uint8_t *fresh) {
    return guard([&] {
        if (!L || (n && (!configs || !names))) throw Error(D4_EINVAL, "null argument");
        for (size_t k = 0; k < n; k++) {
            const bool f = L->log(configs[k], names[k]);
            if (fresh) fresh[k] = f;
        }
    });
}

int d4_ledger_after_run(d4_ledger *L, uint64_t config, uint64_t shape, uint64_t count, uint64_t ms) {
    return guard([&] {
        if (!L) throw Error(D4_EINVAL, "null argument");
        if (config >= L->n_configs) throw Error(D4_EINVAL, "config out of range");
        if (L->cfgs) {
            L->cfgs[config].shape = shape;
            L->cfgs[config].count = count;
            L->cfgs[config].solved = std::max<uint64_t>(ms, 1);
        }
        L->configs_done++;
    });
}

int d4_ledger_progress(const d4_ledger *L, uint64_t *work, uint64_t *configs_done) {
    return guard([&] {
        if (!L) throw Error(D4_EINVAL, "null argument");
        if (work) *work = L->work;
        if (configs_done) *configs_done = L->configs_done;
    });
}

int d4_ledger_names(const d4_ledger *L, uint64_t *out, size_t cap) {
    return guard([&] {
        if (!L || (cap && !out)) throw Error(D4_EINVAL, "null argument");
        const uint64_t *src = L->kind == D4_LEDGER_COUNT ? L->counts.data() : L->nmes;
        if (L->kind == D4_LEDGER_VERIFY) throw Error(D4_ESTATE, "a Verify ledger keeps no names");
        const size_t n = (size_t)std::min<uint64_t>(cap, L->n_names);
        if (n) std::memcpy(out, src, 8 * n);
    });
}

}  // extern "C"
