"""The oracle's and the product's shape geometry against the reference's own
code: src/Shape.cpp, src/Group.hpp and the known-shape tables
(src/known1.cpp, src/known2.cpp), compiled by `make -C oracle ref` into
oracle/_ref/libref.so behind a C shim (oracle/ref_shim.cpp).  Skipped where
that library was not built (it needs the reference's sources)."""
import ctypes
import os
import random

import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")

# SymmetryGroup values (src/Group.hpp:8-19)
GROUPS = [0b00000001, 0b00001001, 0b01101001, 0b00000011, 0b00000101, 0b00010001,
          0b10000001, 0b00001111, 0b10011001, 0b11111111]


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(REF)
    L.ref_transforms.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
    for name in ("ref_canonical",):
        getattr(L, name).restype = ctypes.c_uint64
        getattr(L, name).argtypes = [ctypes.c_uint64, ctypes.c_uint]
    for name in ("ref_normalize", "ref_extend1", "ref_sml_island"):
        getattr(L, name).restype = ctypes.c_uint64
        getattr(L, name).argtypes = [ctypes.c_uint64]
    L.ref_symmetry.restype = ctypes.c_uint
    L.ref_symmetry.argtypes = [ctypes.c_uint64]
    L.ref_connected.restype = ctypes.c_int
    L.ref_connected.argtypes = [ctypes.c_uint64]
    L.ref_known.restype = ctypes.c_size_t
    L.ref_known.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_void_p, ctypes.c_size_t]
    return L


def transforms(ref, s):
    out = (ctypes.c_uint64 * 8)()
    ref.ref_transforms(s, 1, out)
    return list(out)


def known(ref, one_sided, n):
    buf = (ctypes.c_uint64 * 1024)()
    k = ref.ref_known(one_sided, n, buf, 1024)
    return list(buf[:k])


def sample_shapes():
    """Every free polyomino up to 6 cells in every orientation, plus random
    (often disconnected) 8x8 masks, normalised and not."""
    import d4pieces as d4
    out = []
    for n in range(1, 7):
        for s in d4.free_polyominoes(n):
            out += [d4.transform8(s, t) for t in range(8)]
    rng = random.Random(17)
    for _ in range(400):
        v = rng.getrandbits(64) & rng.getrandbits(64) & rng.getrandbits(64)
        out += [v, d4.normalize8(v)] if v else []
    return out


def test_transform_order_matches_reference(ref, oracle):
    import d4pieces as d4
    for s in sample_shapes():
        want = transforms(ref, s)
        assert [oracle.transform(s, t) for t in range(8)] == want, hex(s)
        assert [d4.transform8(s, t) for t in range(8)] == want, hex(s)


def test_canonical_forms_match_reference(ref, oracle):
    for s in sample_shapes():
        for forms in GROUPS:
            assert oracle.canonical(s, forms) == ref.ref_canonical(s, forms), (hex(s), forms)


def test_normalize_matches_reference(ref):
    import d4pieces as d4
    for s in sample_shapes():
        assert d4.normalize8(s) == ref.ref_normalize(s)


def test_unique_orientations_match_symmetry(ref, oracle):
    """Piece's unique transforms (src/Piece.cpp:12-22) number 8 / |stabiliser|,
    the stabiliser being Shape::symmetry()."""
    import d4pieces as d4
    for n in range(1, 8):
        for s in d4.free_polyominoes(n):
            stab = bin(ref.ref_symmetry(s)).count("1")
            assert len(oracle.orients(s, 0xFF)) * stab == 8


def test_polyomino_tables_match_reference(ref, oracle):
    """The product's generated tables and the oracle's enumeration against the
    reference's compiled known-shape tables."""
    import d4pieces as d4
    for n in range(1, 9):
        assert d4.polyominoes(n, 0xFF) == known(ref, 0, n) == oracle.enum_free(n)
        assert d4.polyominoes(n, 0x69) == known(ref, 1, n)


def test_regions_match_reference_islands(ref):
    """workloads.regions (the board filter of the seeded workloads) against
    Shape::connected and Shape::sml_island."""
    import workloads
    rng = random.Random(23)
    for _ in range(500):
        rows = [rng.getrandbits(8) | rng.getrandbits(8) for _ in range(8)]
        v = sum(r << (8 * i) for i, r in enumerate(rows))
        areas = workloads.regions(rows)
        assert sum(areas) == bin(v).count("1")
        assert bool(ref.ref_connected(v)) == (len(areas) <= 1)
        if v:
            assert min(areas) == bin(ref.ref_sml_island(v)).count("1")


def test_orientations_under_subgroups(ref, oracle):
    """The oracle keeps |G| / |G and stabiliser| orientations of a piece under
    each named symmetry group G (orbit-stabiliser), the stabiliser being the
    reference's Shape::symmetry()."""
    import d4pieces as d4
    for n in range(1, 7):
        for s in d4.free_polyominoes(n):
            stab = ref.ref_symmetry(s)
            for g in GROUPS:
                k = len(oracle.orients(s, g))
                assert k * bin(g & stab).count("1") == bin(g).count("1"), (hex(s), g)


def test_embind_shape_surface_matches_reference(ref):
    """pieces_api.Shape8 (the embind Shape8 surface, src/wasm.cpp:48-74)
    against the reference's compiled Shape<8>."""
    import pieces_api as pa
    ref.ref_geometry.argtypes = [ctypes.c_uint64, ctypes.c_void_p]
    for s in sample_shapes():
        if not s:
            continue
        sh = pa.Shape8(s)
        g = (ctypes.c_uint64 * 6)()
        ref.ref_geometry(s, g)
        assert [sh.left, sh.width, sh.right, sh.top, sh.height, sh.bottom] == list(g), hex(s)
        raw = (ctypes.c_uint64 * 8)()
        ref.ref_transforms(s, 0, raw)
        fns = [sh.transform0, sh.transform1, sh.transform2, sh.transform3, sh.transform4,
               sh.transform5, sh.transform6, sh.transform7]
        assert [f(False).value for f in fns] == list(raw)
        assert [f(True).value for f in fns] == transforms(ref, s)
        assert sh.symmetry == ref.ref_symmetry(s)
        assert int(sh.classify) == ref.ref_symmetry(s)
        assert sh.normalize().value == ref.ref_normalize(s)
        for forms in GROUPS:
            assert sh.canonical_form(forms).value == ref.ref_canonical(s, forms)


def test_embind_piece_and_groups(ref, oracle):
    """Piece8 placements (normal, enabled, duplicate) as the Piece constructor
    (src/Piece.cpp:12-22) makes them, compared with the oracle's unique
    orientations; groupProduct against the reference's rule table where that
    gives a group, else the group closure of the reference's mask."""
    import d4pieces as d4
    import pieces_api as pa
    ref.ref_group_product.restype = ctypes.c_uint
    ref.ref_group_product.argtypes = [ctypes.c_uint, ctypes.c_uint]
    for n in range(1, 7):
        for s in d4.free_polyominoes(n):
            for sym in GROUPS:
                p = pa.Piece8(pa.Shape8(s), sym)
                on = [(pl.normal.value, t) for t, pl in enumerate(p.placements) if pl.enabled]
                assert on == oracle.orients(s, sym)
                assert all(pl.duplicate != pl.enabled for pl in p.placements)
                assert len(p.placements) == bin(sym).count("1") and p.count == 1
    for a in GROUPS:
        for b in GROUPS:
            r = ref.ref_group_product(a, b)
            got = pa.groupProduct(a, b)
            assert got in GROUPS
            assert got == r if r in GROUPS else got == pa.groupProduct(r, r)
            assert pa.subgroup(got, a) and pa.subgroup(got, b)
    assert pa.groupProduct(pa.SymmetryGroup.C4, pa.SymmetryGroup.D1_P) == pa.SymmetryGroup.D4
    assert [pa.shape_count(n) for n in range(1, 9)] == [1, 1, 2, 5, 12, 35, 108, 369]
    assert pa.shape_at8(5, 0).value == known(ref, 0, 5)[0]


def _ref11(ref):
    u64, vp = ctypes.c_uint64, ctypes.c_void_p
    ref.ref11_transforms.argtypes = [u64, u64, ctypes.c_int, vp]
    ref.ref11_geometry.argtypes = [u64, u64, vp]
    ref.ref11_canonical.argtypes = [u64, u64, ctypes.c_uint, vp]
    ref.ref11_symmetry.restype = ctypes.c_uint
    ref.ref11_symmetry.argtypes = [u64, u64]
    ref.ref11_normalize.argtypes = [u64, u64, vp]
    ref.ref11_shape_at.argtypes = [ctypes.c_uint, ctypes.c_size_t, vp]
    return ref


def _split(v):
    return v & ((1 << 64) - 1), v >> 64


def _join(buf, i=0):
    return buf[2 * i] | buf[2 * i + 1] << 64


def sample_shapes11():
    """Free polyominoes up to 6 cells placed at random offsets in the 11x11
    frame, shapes touching the right and bottom edges, full rows and
    columns, and random (often disconnected) 121-bit masks."""
    import pieces_api as pa
    rng = random.Random(11)
    out = [1, (1 << 11) - 1, sum(1 << (11 * i) for i in range(11)), (1 << 121) - 1,
           1 << 120, 1 << 10, 1 << 110]
    for n in range(1, 7):
        for i in range(pa.shape_count(n)):
            s = pa.shape_at11(n, i)
            for t in range(8):
                v = s._t(t, True).value
                out.append(v)
                out.append(v << (11 * rng.randint(0, 5) + rng.randint(0, 5)) & ((1 << 121) - 1))
    for _ in range(300):
        out.append(rng.getrandbits(121) & rng.getrandbits(121) & rng.getrandbits(121))
    return [v for v in out if v]


def test_shape11_against_reference(ref):
    """Shape11's transforms (normalised and in place), normalize, margins and
    extents (including the whole-word shifts of a shape spanning every
    column), canonical forms and symmetry equal the compiled Shape<11>."""
    import pieces_api as pa
    ref = _ref11(ref)
    buf = (ctypes.c_uint64 * 16)()
    g = (ctypes.c_uint64 * 6)()
    two = (ctypes.c_uint64 * 2)()
    for v in sample_shapes11():
        lo, hi = _split(v)
        sh = pa.Shape11(v)
        for norm in (1, 0):
            ref.ref11_transforms(lo, hi, norm, buf)
            assert [sh._t(t, bool(norm)).value for t in range(8)] == \
                [_join(buf, t) for t in range(8)], (hex(v), norm)
        ref.ref11_normalize(lo, hi, two)
        assert sh.normalize().value == _join(two)
        ref.ref11_geometry(lo, hi, g)
        assert [sh.left, sh.width, sh.right, sh.top, sh.height, sh.bottom] == list(g), hex(v)
        assert sh.symmetry == ref.ref11_symmetry(lo, hi)
        for forms in (0, 1, 0x0F, 0x69, 0xFF, 0x90):
            ref.ref11_canonical(lo, hi, forms, two)
            assert sh.canonical_form(forms).value == _join(two), (hex(v), forms)


def test_shape_at11_against_reference(ref):
    """shape_at11 equals shape_at<11> for every free polyomino of 1-8 cells."""
    import pieces_api as pa
    ref = _ref11(ref)
    two = (ctypes.c_uint64 * 2)()
    for n in range(1, 9):
        for i in range(pa.shape_count(n)):
            ref.ref11_shape_at(n, i, two)
            assert pa.shape_at11(n, i).value == _join(two), (n, i)
