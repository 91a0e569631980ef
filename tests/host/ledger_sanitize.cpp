// Host-only checks of the naming, ledger and polyomino code (d4_ledger.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitize.py: name/resolve round trips (src/naming.inl:6-72),
// out-of-range ids and names, ledgers of every kind over temporary files
// (src/search.cpp:25-226) including reopen and refused sizes, and the
// polyomino tables at every size and symmetry.  Exit status 0 = all passed.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <unistd.h>
#include "d4.h"

static int fails = 0;
#define CHECK(x) do { if (!(x)) { fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #x); fails++; } } while (0)

static void naming_case(uint64_t m, uint64_t M, uint64_t n, uint64_t N, uint64_t t, std::vector<uint64_t> arr) {
    d4_naming *nm = nullptr;
    int rc = d4_naming_create(m, M, n, N, t, arr.data(), &nm);
    if (rc) { printf("naming rc %d for (%lu..%lu, %lu..%lu, %lu)\n", rc, m, M, n, N, t); return; }
    const uint64_t S = d4_naming_size(nm), P = d4_naming_size_pieces(nm);
    std::vector<uint8_t> used(P + 1);
    uint64_t lim = S < 20000 ? S : 20000;
    for (uint64_t k = 0; k < lim; k++) {
        uint64_t idx = S < 20000 ? k : (k * 2654435761ull) % S;
        CHECK(d4_naming_resolve(nm, idx, used.data()) == 0);
        uint64_t back = 0;
        CHECK(d4_naming_name(nm, used.data(), &back) == 0);
        CHECK(back == idx);
    }
    CHECK(d4_naming_resolve(nm, S, used.data()) != 0);
    for (uint64_t id = 0; id <= P; id++) {
        uint64_t mm = 0, ii = 0;
        int r = d4_naming_resolve_piece(nm, id, &mm, &ii);
        if (id < P) { CHECK(r == 0); CHECK(d4_naming_name_piece(nm, mm, ii) == id); }
        else CHECK(r != 0);
    }
    CHECK(d4_naming_name_piece(nm, M + 1, 0) == D4_NO_NAME);
    CHECK(d4_naming_name_piece(nm, 0, 0) == D4_NO_NAME || m == 0);
    // random sets, most not named
    srand(7);
    for (int k = 0; k < 2000; k++) {
        for (uint64_t id = 0; id < P; id++) used[id] = rand() % 4 == 0;
        uint64_t nme = 0;
        CHECK(d4_naming_name(nm, used.data(), &nme) == 0);
        if (nme != D4_NO_NAME) {
            std::vector<uint8_t> u2(P);
            CHECK(d4_naming_resolve(nm, nme, u2.data()) == 0);
            CHECK(memcmp(u2.data(), used.data(), P) == 0);
        }
    }
    printf("naming (%lu..%lu, %lu..%lu, %lu): size %lu pieces %lu\n", m, M, n, N, t, S, P);
    d4_naming_destroy(nm);
}

static void ledger_case(int kind, uint64_t C, uint64_t Nn) {
    char path[] = "/tmp/d4ledgerXXXXXX";
    int fd = mkstemp(path);
    close(fd);
    d4_ledger *L = nullptr;
    CHECK(d4_ledger_open(kind, kind >= 2 ? path : nullptr, C, Nn, &L) == 0);
    if (!L) return;
    std::vector<uint64_t> cf, nmv;
    for (uint64_t c = 0; c < C; c++) {
        int run = 0;
        CHECK(d4_ledger_should_run(L, c, 0x1234 + c, &run) == 0);
        for (uint64_t k = 0; k < 5; k++) {
            int fresh = 0;
            CHECK(d4_ledger_log(L, c, (c * 7 + k * 3) % Nn, &fresh) == 0);
            cf.push_back(c); nmv.push_back((c + k) % Nn);
        }
        int fresh = 0;
        CHECK(d4_ledger_log(L, c, D4_NO_NAME, &fresh) == 0);
        CHECK(fresh == 0);
        CHECK(d4_ledger_after_run(L, c, 0x1234 + c, 5, 0) == 0);
    }
    std::vector<uint8_t> fr(cf.size());
    CHECK(d4_ledger_log_many(L, cf.data(), nmv.data(), cf.size(), fr.data()) == 0);
    CHECK(d4_ledger_log_many(L, cf.data(), nmv.data(), cf.size(), nullptr) == 0);
    uint64_t w = 0, done = 0;
    CHECK(d4_ledger_progress(L, &w, &done) == 0);
    std::vector<uint64_t> names(Nn + 3);
    int r = d4_ledger_names(L, names.data(), Nn + 3);
    CHECK(kind == 0 ? r != 0 : r == 0);
    // out-of-range config / name
    int fresh = 0;
    CHECK(d4_ledger_log(L, C, 0, &fresh) != 0);
    CHECK(d4_ledger_log(L, 0, Nn, &fresh) != 0);
    d4_ledger_close(L);
    if (kind >= 2) {
        // reopen: same sizes accepted, other sizes refused
        CHECK(d4_ledger_open(kind, path, C, Nn, &L) == 0);
        int run = 1;
        CHECK(d4_ledger_should_run(L, 0, 0x1234, &run) == 0);
        CHECK(run == 0);
        d4_ledger_close(L);
        CHECK(d4_ledger_open(kind, path, C + 1, Nn, &L) != 0);
        CHECK(d4_ledger_open(kind, path, C, Nn - 1, &L) != 0);
    }
    unlink(path);
    printf("ledger kind %d (%lu configs, %lu names): work %lu done %lu\n", kind, C, Nn, w, done);
}

int main() {
    naming_case(4, 6, 2, 4, 20, {0, 0, 0, 0, 5, 12, 35});
    naming_case(1, 5, 1, 8, 15, {0, 1, 1, 2, 5, 12});
    naming_case(3, 6, 1, 14, 40, {0, 0, 0, 2, 5, 12, 35});
    naming_case(5, 5, 12, 12, 60, {0, 0, 0, 0, 0, 12});
    naming_case(2, 8, 1, 30, 100, {0, 0, 1, 2, 5, 12, 35, 108, 369});
    for (int kind = 0; kind < 4; kind++) {
        ledger_case(kind, 13, 29);
        ledger_case(kind, 64, 8);
        ledger_case(kind, 1, 1);
    }
    for (unsigned n = 0; n <= 9; n++)
        for (uint32_t sym : {0xFFu, 0x69u, 0x01u}) {
            size_t tot = d4_polyominoes(n, sym, nullptr, 0);
            std::vector<uint64_t> v(tot + 1);
            CHECK(d4_polyominoes(n, sym, v.data(), tot) == tot);
            CHECK(d4_polyominoes(n, sym, v.data(), tot / 2) == tot);
            printf("polyominoes n=%u sym=%#x: %zu\n", n, sym, tot);
        }
    printf("fails %d\n", fails);
    return fails != 0;
}
