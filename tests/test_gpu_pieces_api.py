"""solve8 of the embind surface mirror (src/wasm.cpp:35, src/Piece.cpp:65-102)
on the device, against the oracle's solve with the same per-placement masks:
every solution's steps (piece, transform id, translation, matrix, placed
shape) and cell map, in the CPU searcher's order."""
import pytest

import d4pieces as d4
import pieces_api as pa

pytestmark = pytest.mark.gpu


def shape_of(rows):
    return sum(r << (8 * i) for i, r in enumerate(rows))


def rect(w, h, blocked=()):
    rows = [(1 << w) - 1] * h
    for r, c in blocked:
        rows[r] &= 0xFF ^ (1 << c)
    return rows


def check(oracle, pieces, rows, single=False, min_solutions=1):
    print(f"solve8 case: {len(pieces)} pieces, rows {rows}, single {single}", flush=True)
    got = pa.solve8(pieces, pa.Shape8(shape_of(rows)), single)
    shapes = [d4.normalize8(p.shape.value) for p in pieces]
    en = [sum(1 << i for i, pl in enumerate(p.placements) if pl.enabled) for p in pieces]
    want = oracle.solve(shapes, [p.count for p in pieces], rows, single=single,
                        sym=pieces[0].sym, enabled=en)
    assert len(got) == len(want) >= min_solutions
    sym_t = [t for t in range(8) if pieces[0].sym >> t & 1]
    for sol, raw in zip(got, want):
        assert [(st.piece_id, st.trs_id, st.y, st.x) for st in sol.steps] == list(raw)
        cover = 0
        for st, (pid, trs, y, x) in zip(sol.steps, raw):
            # the placed shape is the transform's; the matrix follows the
            # placement index, as make_step does
            assert st.shape == d4.transform8(shapes[pid], sym_t[trs]) << (8 * y + x)
            abcd = (ctypes_matrix(oracle, trs))
            assert (st.a, st.b, st.c, st.d) == abcd
            assert not cover & st.shape
            cover |= st.shape
        assert cover == shape_of(rows)
        assert len(sol.map) == 8 and all(len(r) == 8 for r in sol.map)
        for r in range(8):
            for c in range(8):
                k = sol.map[r][c]
                assert (k >= 0) == bool(cover >> (8 * r + c) & 1)
                if k >= 0:
                    assert sol.steps[k].shape >> (8 * r + c) & 1
    return got


def ctypes_matrix(oracle, trs):
    import ctypes
    m = (ctypes.c_int * 4)()
    oracle.L.or_step_matrix(trs, m)
    return tuple(m)


def test_solve8_default_pieces(gpu, oracle):
    p5 = [pa.Piece8(pa.shape_at8(5, i)) for i in range(pa.shape_count(5))]
    scott = rect(8, 8, [(3, 3), (3, 4), (4, 3), (4, 4)])
    got = check(oracle, p5, scott, min_solutions=520)
    assert len(got) == 520
    first = check(oracle, p5, scott, single=True)
    assert len(first) == 1
    assert [(s.piece_id, s.trs_id, s.x, s.y) for s in first[0].steps] == \
        [(s.piece_id, s.trs_id, s.x, s.y) for s in got[0].steps]


def test_solve8_custom_placements(gpu, oracle):
    """Placements switched off and on the way the UI does (Placement8.enabled):
    the X pentomino's duplicates on (8 solutions per tiling), the L
    tetromino's reflections off, counts above 1."""
    p5 = [pa.Piece8(pa.shape_at8(5, i)) for i in range(12)]
    x = next(i for i, p in enumerate(p5) if p.shape.value == 0x020702)
    for pl in p5[x].placements:
        pl.enabled = True
    check(oracle, p5, rect(8, 8, [(3, 3), (3, 4), (4, 3), (4, 4)]), min_solutions=8 * 520)
    p4 = [pa.Piece8(pa.shape_at8(4, i)) for i in range(5)]
    for p in p4:
        p.count = 2
    for t in (1, 2, 4, 7):  # reflections of every tetromino off
        for p in p4:
            p.placements[t].enabled = False
    check(oracle, p4, rect(5, 8), min_solutions=1)
    check(oracle, p4, rect(5, 8), single=True)


def test_solve8_symmetry_subgroup_and_edges(gpu, oracle):
    """Pieces built under the C4 mask (transform ids count the rotations
    only), an unsolvable board and the empty board."""
    p4 = [pa.Piece8(pa.shape_at8(4, i), pa.SymmetryGroup.C4) for i in range(5)]
    for p in p4:
        p.count = 2
    check(oracle, p4, rect(4, 5), min_solutions=1)
    p5 = [pa.Piece8(pa.shape_at8(5, i)) for i in range(12)]
    assert pa.solve8(p5, pa.Shape8(shape_of(rect(4, 1))), False) == []
    empty = pa.solve8(p5, pa.Shape8(0), False)
    assert len(empty) == 1 and empty[0].steps == []
    assert all(v == -1 for row in empty[0].map for v in row)
