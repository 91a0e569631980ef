// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY: a C entry point per reference
// function the tests compare against.  Built by `make -C oracle ref` together
// with the reference's own src/Shape.cpp, src/Group.cpp, src/known1.cpp and
// src/known2.cpp (compiled where they lie, under /root/reference) into
// oracle/_ref/libref.so.  Nothing here restates the reference: every call
// goes to its code.  Only tests/ load the library.
#include <cstddef>
#include <cstdint>

#include "Group.hpp"
#include "Shape.hpp"
#include "known.hpp"

extern "C" {

// Shape<8>::transforms(norm) (src/Shape.cpp:28-40): the 8 D4 images in the
// reference's transform order.
void ref_transforms(uint64_t s, int norm, uint64_t *out) {
    const auto t = Shape<8>{s}.transforms(norm != 0);
    for (int i = 0; i < 8; i++) out[i] = t[i].get_value();
}

// Shape<8>::left / width / right / top / height / bottom (src/Shape.hpp:126-168)
void ref_geometry(uint64_t s, uint64_t *out) {
    const Shape<8> sh{s};
    out[0] = sh.left();
    out[1] = sh.width();
    out[2] = sh.right();
    out[3] = sh.top();
    out[4] = sh.height();
    out[5] = sh.bottom();
}

// Shape<8>::canonical_form(forms) (src/Shape.cpp:57-66)
uint64_t ref_canonical(uint64_t s, unsigned forms) { return Shape<8>{s}.canonical_form(forms).get_value(); }

// Shape<8>::symmetry() (src/Shape.cpp:68-79)
unsigned ref_symmetry(uint64_t s) { return Shape<8>{s}.symmetry(); }

// Shape<8>::normalize() (src/Shape.hpp:182-191)
uint64_t ref_normalize(uint64_t s) { return Shape<8>{s}.normalize().get_value(); }

// Shape<8>::extend1 / connected / sml_island (src/Shape.cpp:81-122)
uint64_t ref_extend1(uint64_t s) { return Shape<8>{s}.extend1().get_value(); }
int ref_connected(uint64_t s) { return Shape<8>{s}.connected() ? 1 : 0; }
uint64_t ref_sml_island(uint64_t s) { return Shape<8>{s}.sml_island().get_value(); }

// SymmetryGroup product (src/Group.hpp:27-50)
unsigned ref_group_product(unsigned a, unsigned b) {
    return static_cast<unsigned>(static_cast<SymmetryGroup>(a) * static_cast<SymmetryGroup>(b));
}

// known_shapes / known_C_shapes (src/known1.cpp, src/known2.cpp): n-cell
// table, min(total, cap) values written, total returned.
size_t ref_known(int one_sided, unsigned n, uint64_t *out, size_t cap) {
    if (n < 1 || n > 8) return 0;
    const uint64_t *const *tab = one_sided ? known_C_shapes : known_shapes;
    const size_t *cnt = one_sided ? shapes_C_count : shapes_count;
    for (size_t i = 0; i < cnt[n] && i < cap; i++) out[i] = tab[n][i];
    return cnt[n];
}

// Shape<11> (121 bits, two 64-bit halves in and out): the same functions
// on the reference's 11x11 frame (src/Shape.cpp:158-176 instantiates them).
static Shape<11> s11(uint64_t lo, uint64_t hi) {
    return Shape<11>{(static_cast<__uint128_t>(hi) << 64) | lo};
}
static void put11(Shape<11> s, uint64_t *out) {
    const __uint128_t v = s.get_value();
    out[0] = static_cast<uint64_t>(v);
    out[1] = static_cast<uint64_t>(v >> 64);
}

void ref11_transforms(uint64_t lo, uint64_t hi, int norm, uint64_t *out) {
    const auto t = s11(lo, hi).transforms(norm != 0);
    for (int i = 0; i < 8; i++) put11(t[i], out + 2 * i);
}

void ref11_geometry(uint64_t lo, uint64_t hi, uint64_t *out) {
    const Shape<11> sh = s11(lo, hi);
    out[0] = sh.left();
    out[1] = sh.width();
    out[2] = sh.right();
    out[3] = sh.top();
    out[4] = sh.height();
    out[5] = sh.bottom();
}

void ref11_canonical(uint64_t lo, uint64_t hi, unsigned forms, uint64_t *out) {
    put11(s11(lo, hi).canonical_form(forms), out);
}

unsigned ref11_symmetry(uint64_t lo, uint64_t hi) { return s11(lo, hi).symmetry(); }

void ref11_normalize(uint64_t lo, uint64_t hi, uint64_t *out) { put11(s11(lo, hi).normalize(), out); }

// shape_at<11>(n, i) (src/known.hpp:15-18)
void ref11_shape_at(unsigned n, size_t i, uint64_t *out) { put11(shape_at<11>(n, i), out); }

}  // extern "C"
