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


def _solve_l():
    import os
    import sys
    from conftest import ROOT
    p = os.path.join(ROOT, "oracle")
    if p not in sys.path:
        sys.path.insert(0, p)
    import solve_l
    return solve_l


def board11(cells):
    return pa.Shape11(sum(1 << (11 * r + c) for r, c in cells))


def check11(pieces, board, single=False, min_solutions=1):
    """solve11 on the device equals the L = 11 restatement (oracle/solve_l.py):
    the same solutions in the same order, each step's piece, placement index,
    make_step matrix, translation and placed shape, and the 11 x 11 map."""
    S = _solve_l()
    lib = []
    for p in pieces:
        q = S.Piece(p.shape.normalize().value, 11, p.sym, p.count)
        for i, pl in enumerate(p.placements):
            q.placements[i][2] = pl.enabled
        lib.append(q)
    expect = S.solve(lib, board.value, 11, single=single)
    got = pa.solve11(pieces, board, single=single)
    assert len(got) == len(expect) >= min(min_solutions, len(expect))
    assert len(expect) >= min_solutions
    for sol, ex in zip(got, expect):
        assert [(st.piece_id, st.trs_id, st.x, st.y, st.shape.value) for st in sol.steps] == \
            [tuple(e) for e in ex]
        for st in sol.steps:
            assert (st.a, st.b, st.c, st.d) == S.STEP_MATRIX[st.trs_id]
        for r in range(11):
            for c in range(11):
                cover = [i for i, st in enumerate(sol.steps) if st.shape.test(r, c)]
                assert sol.map[r][c] == (cover[0] if cover else -1)
    return got


def test_solve11_board_within_8_columns(gpu):
    """A board in columns 3-10 (searched as it lies, shifted): the five
    tetrominoes and the L tromino x2 on 8x3 with a notch; all solutions and
    the first."""
    pieces = [pa.Piece11(pa.shape_at11(4, i)) for i in range(5)] + [pa.Piece11(pa.shape_at11(3, 1))]
    pieces[-1].count = 2
    cells = [(r, c) for r in range(3) for c in range(3, 11) if (r, c) not in ((0, 3),)]
    b = board11(cells)
    check11(pieces, b, min_solutions=40)
    check11(pieces, b, single=True)


def test_solve11_wide_board_transposed(gpu):
    """Boards 9-11 columns wide and at most 8 rows tall are searched
    transposed; the solutions come back in the CPU searcher's order: dominoes
    on 2x11 (144), straight trominoes on 3x11 (41), a mixed set on 9x3 with a
    notch under the D4, C4, D1_X and identity-only symmetry masks (5,748 /
    5,748 / 96 / 6 tilings), and the O tetromino with its duplicate
    placements switched on."""
    dom = pa.Piece11(pa.shape_at11(2, 0))
    dom.count = 11
    got = check11([dom], board11([(r, c) for r in range(2) for c in range(11)]), min_solutions=144)
    assert len(got) == 144
    tro = pa.Piece11(pa.shape_at11(3, 0))
    tro.count = 11
    assert len(check11([tro], board11([(r, c) for r in range(3) for c in range(11)]))) == 41
    for sym in (0xFF, 0x69, 0x03, 0x01):
        pieces = [pa.Piece11(pa.shape_at11(3, i), sym) for i in range(2)] + \
            [pa.Piece11(pa.shape_at11(2, 0), sym), pa.Piece11(pa.shape_at11(4, 3), sym)]
        for p, c in zip(pieces, (3, 3, 2, 1)):
            p.count = c
        cells = [(r, c) for r in range(3) for c in range(9) if (r, c) not in ((2, 8),)]
        check11(pieces, board11(cells), min_solutions=1)
        check11(pieces, board11(cells), single=True)
    # duplicates switched on: the O tetromino's 8 placements, all enabled
    sq = pa.Piece11(pa.shape_at11(4, 3))
    for pl in sq.placements:
        pl.enabled = True
    sq.count = 2
    ln = pa.Piece11(pa.shape_at11(2, 0))
    ln.count = 6
    cells = [(r, c) for r in range(2) for c in range(10)]
    got = check11([sq, ln], board11(cells), min_solutions=8)
    check11([sq, ln], board11(cells), single=True)


def test_solve11_refusals_and_edges(gpu):
    """A board spanning more than 8 rows and 8 columns is refused; the empty
    board has one empty solution; an untileable board has none."""
    p = pa.Piece11(pa.shape_at11(2, 0))
    p.count = 60
    with pytest.raises(ValueError):
        pa.solve11([p], board11([(r, c) for r in range(9) for c in range(9)]))
    sols = pa.solve11([p], pa.Shape11(0))
    assert len(sols) == 1 and sols[0].steps == [] and all(v == -1 for r in sols[0].map for v in r)
    assert pa.solve11([p], board11([(0, 0), (0, 2), (5, 10)])) == []
