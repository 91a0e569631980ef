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

}  // extern "C"
