// Host-only checks of the row-step guards in d4_host.h: a row whose unique
// states reach 2^32 is refused before the compaction (whose offsets are
// 32-bit) runs, and the compaction check compares its counts in 64 bits, so
// counts that agree only modulo 2^32 are an error.  Exit status 0 = passed.
#include <cstdio>
#include "d4_host.h"

static int fails = 0;
#define CHECK(x) do { if (!(x)) { fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #x); fails++; } } while (0)

template <class F>
static int code_of(F &&f) {
    try {
        f();
        return 0;
    } catch (const d4::Error &e) {
        return e.code;
    }
}

int main() {
    CHECK(code_of([] { d4::check_row_states(0); }) == 0);
    CHECK(code_of([] { d4::check_row_states(0xffffffffull); }) == 0);
    CHECK(code_of([] { d4::check_row_states(0x100000000ull); }) == D4_ERANGE);
    CHECK(code_of([] { d4::check_row_states(3110000000ull * 2); }) == D4_ERANGE);
    CHECK(code_of([] { d4::check_compacted(7, 7); }) == 0);
    CHECK(code_of([] { d4::check_compacted(0x100000007ull, 7); }) == D4_EDEVICE);
    CHECK(code_of([] { d4::check_compacted(7, 0x100000007ull); }) == D4_EDEVICE);
    CHECK(code_of([] { d4::check_compacted(6, 7); }) == D4_EDEVICE);
    printf("fails %d\n", fails);
    return fails != 0;
}
