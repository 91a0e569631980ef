"""Pins the oracle (CPU restatement of src/Piece.cpp solve/solve_count and the
Shape geometry) against the reference's fixtures, closed forms and brute
force -- before it is trusted as the GPU parity checker."""
import itertools
import json
import os
import random

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_shapes.json")

# OEIS A000105 (free) / A001168 (fixed) polyominoes, n = 1..8
A000105 = [1, 1, 2, 5, 12, 35, 108, 369]
A001168 = [1, 2, 6, 19, 63, 216, 760, 2725]


def rect(w, h, blocked=()):
    rows = [(1 << w) - 1] * h
    for r, c in blocked:
        rows[r] &= 0xFF ^ (1 << c)
    return rows


@pytest.fixture(scope="module")
def golden():
    return json.load(open(GOLDEN))


@pytest.fixture(scope="module")
def p5(oracle):
    return oracle.enum_free(5)


def test_enum_matches_reference_tables(oracle, golden):
    assert golden["free_counts"][1:] == A000105
    for n in range(1, 9):
        got = oracle.enum_free(n)
        assert len(got) == A000105[n - 1]
        assert got == golden["free"][str(n)], n


def test_transform_group(oracle):
    # every transform of a normalised shape is normalised; involutions square
    # to the identity and the two quarter turns are inverses (Group.cpp:3-32)
    rng = random.Random(1)
    shapes = oracle.enum_free(6) + oracle.enum_free(7)
    inv = {0: 0, 1: 1, 2: 2, 3: 3, 4: 4, 7: 7, 5: 6, 6: 5}
    for s in rng.sample(shapes, 40):
        assert oracle.transform(s, 0) == s
        for t in range(8):
            assert oracle.transform(oracle.transform(s, t), inv[t]) == s
        # closure: transforming any transform stays within the orbit
        orbit = {oracle.transform(s, t) for t in range(8)}
        for a in orbit:
            assert {oracle.transform(a, t) for t in range(8)} == orbit


def test_transform_concrete(oracle):
    # L tetromino 0b1_00000111 (row0: ###, row1: #..)
    L = 0b1_0000_0111 | 0
    L = 7 | (1 << 8)
    assert oracle.transform(L, 1) == 7 | (1 << 10)            # flip X: ###/..#
    assert oracle.transform(L, 2) == (1 << 0) | (7 << 8)      # flip Y: #../###
    assert oracle.transform(L, 4) == 0b11 | (1 << 8) | (1 << 16)  # transpose


def test_fixed_orientation_counts(oracle):
    for n in range(1, 8):
        total = sum(len(oracle.orients(s)) for s in oracle.enum_free(n))
        assert total == A001168[n - 1], n


def test_pentomino_symmetry_classes(oracle, p5):
    # X has 1 orientation, I 2, T/U/V/W/Z 4, F/L/N/P/Y 8 -> 63 in total
    sizes = sorted(len(oracle.orients(s)) for s in p5)
    assert sizes == [1, 2, 4, 4, 4, 4, 4, 8, 8, 8, 8, 8]


@pytest.mark.parametrize("w,h,expect", [(6, 10, 9356), (5, 12, 4040), (4, 15, 1472),
                                        (3, 20, 8)])
def test_pentomino_rectangles(oracle, p5, w, h, expect):
    assert oracle.count(p5, [1] * 12, rect(w, h)) == expect


def test_scott_centre_hole(oracle, p5):
    assert oracle.count(p5, [1] * 12, rect(8, 8, [(3, 3), (3, 4), (4, 3), (4, 4)])) == 520


def test_domino_fibonacci(oracle):
    fib = [1, 1]
    for _ in range(20):
        fib.append(fib[-1] + fib[-2])
    for n in range(1, 14):
        assert oracle.count([3], [n], rect(2, n)) == fib[n], n


def test_domino_kasteleyn_8x8(oracle):
    assert oracle.count([3], [32], rect(8, 8)) == 12988816


def test_domino_kasteleyn_small(oracle):
    import math
    for m, n in [(3, 4), (4, 4), (4, 5), (5, 6), (6, 6)]:
        prod = 1.0
        for j in range(1, m + 1):
            for k in range(1, n + 1):
                prod *= 4 * math.cos(math.pi * j / (m + 1)) ** 2 + \
                    4 * math.cos(math.pi * k / (n + 1)) ** 2
        want = round(prod ** 0.25) if m * n % 2 == 0 else 0
        assert oracle.count([3], [m * n // 2], rect(m, n)) == want, (m, n)


def test_straight_tromino_recurrence(oracle):
    a = [1, 1, 1]
    for n in range(3, 16):
        a.append(a[n - 1] + a[n - 3])
    for n in range(1, 16):
        assert oracle.count([7], [n], rect(3, n)) == a[n], n


def _placements(shapes, counts, rows, oracle):
    H = len(rows)
    board = {(r, c) for r in range(H) for c in range(8) if (rows[r] >> c) & 1}
    out = []
    for i, s in enumerate(shapes):
        for nm, _ in oracle.orients(s):
            cells = [(b // 8, b % 8) for b in range(64) if (nm >> b) & 1]
            for dy in range(H):
                for dx in range(8):
                    pc = frozenset((r + dy, c + dx) for r, c in cells)
                    if pc <= board:
                        out.append((i, pc))
    return board, out


def _brute(shapes, counts, rows, oracle):
    """Exact covers as sets of placements (order-free), counts respected."""
    board, pls = _placements(shapes, counts, rows, oracle)
    total = 0

    def rec(start, covered, used):
        nonlocal total
        if covered == board:
            total += 1
            return
        for j in range(start, len(pls)):
            i, pc = pls[j]
            if used[i] < counts[i] and not (pc & covered):
                used[i] += 1
                rec(j + 1, covered | pc, used)
                used[i] -= 1

    rec(0, frozenset(), [0] * len(shapes))
    return total


def test_brute_force_tiny(oracle):
    rng = random.Random(7)
    pool = [1, 3, 7, 259] + oracle.enum_free(4)
    checked = 0
    for trial in range(60):
        k = rng.randint(2, 4)
        shapes = sorted(rng.sample(pool, k))
        counts = [rng.randint(1, 3) for _ in shapes]
        w, h = rng.randint(2, 4), rng.randint(2, 4)
        blocked = [(rng.randrange(h), rng.randrange(w)) for _ in range(rng.randint(0, 2))]
        rows = rect(w, h, blocked)
        want = _brute(shapes, counts, rows, oracle)
        assert oracle.count(shapes, counts, rows) == want, (shapes, counts, rows)
        checked += want > 0
    assert checked >= 10


def _cover_ok(sol, shapes, counts, rows, oracle):
    board = {(r, c) for r in range(len(rows)) for c in range(8) if (rows[r] >> c) & 1}
    covered = set()
    used = [0] * len(shapes)
    for pid, trs, y, x in sol:
        nm = dict((t, n) for n, t in oracle.orients(shapes[pid]))[trs]
        cells = {(b // 8 + y, b % 8 + x) for b in range(64) if (nm >> b) & 1}
        assert not (cells & covered)
        covered |= cells
        used[pid] += 1
    return covered == board and all(u <= c for u, c in zip(used, counts))


@pytest.mark.slow
def test_solve_6x10_invariants(oracle, p5):
    rows = rect(6, 10)
    sols = oracle.solve(p5, [1] * 12, rows)
    assert len(sols) == 9356 == len(set(map(tuple, sols)))
    # DFS order: solutions come out sorted by their (piece, trs) choice sequence
    keys = [[(p, t) for p, t, _, _ in s] for s in sols]
    assert keys == sorted(keys)
    first = oracle.solve(p5, [1] * 12, rows, single=True)
    assert first == sols[:1]
    for s in sols[::97]:
        assert _cover_ok(s, p5, [1] * 12, rows, oracle)
    # board symmetry: mirroring every solution left-right maps the set onto
    # itself (as cell partitions labelled by piece)
    def partition(sol, mirror):
        out = []
        for pid, trs, y, x in sol:
            nm = dict((t, n) for n, t in oracle.orients(p5[pid]))[trs]
            cells = frozenset((b // 8 + y, (5 - (b % 8 + x)) if mirror else b % 8 + x)
                              for b in range(64) if (nm >> b) & 1)
            out.append((pid, cells))
        return frozenset(out)
    plain = {partition(s, False) for s in sols}
    assert {partition(s, True) for s in sols} == plain


def test_solve_matches_count_small(oracle):
    rng = random.Random(11)
    pool = [3, 7, 259] + oracle.enum_free(4)
    for _ in range(25):
        shapes = sorted(rng.sample(pool, 3))
        counts = [rng.randint(1, 3) for _ in shapes]
        rows = rect(rng.randint(2, 5), rng.randint(2, 5))
        sols = oracle.solve(shapes, counts, rows)
        assert len(sols) == oracle.count(shapes, counts, rows)
        for s in sols:
            assert _cover_ok(s, shapes, counts, rows, oracle)


def test_subset_tilings(oracle):
    # pieces need not all be used (solve only prunes by area): a 2x2 board
    # with a monomino x4 and a square tetromino x1 has 2 tilings
    assert oracle.count([1, 15 & 0x3 | (0x3 << 8)], [4, 1], rect(2, 2)) == 2


def default_masks(oracle, shapes, sym=0xFF):
    """Piece::Placement::enabled as the Piece constructor sets it: every
    placement that is not a duplicate of an earlier one."""
    return [sum(1 << t for _, t in oracle.orients(s, sym)) for s in shapes]


def test_enabled_placements(oracle, p5):
    """or_solve_ex (per-placement masks, Piece.hpp:16-20 / Piece.inl:8-10):
    the constructor's masks give or_solve's output; enabling the X
    pentomino's 7 duplicate placements multiplies every tiling by 8; keeping
    only the rotations equals the C4 symmetry mask; a piece with no placement
    is a piece left out."""
    scott = rect(8, 8, [(3, 3), (3, 4), (4, 3), (4, 4)])
    base = oracle.solve(p5, [1] * 12, scott)
    assert len(base) == 520
    assert oracle.solve(p5, [1] * 12, scott, enabled=default_masks(oracle, p5)) == base
    x = p5.index(0x020702)  # the X pentomino: all 8 transforms coincide
    assert len(oracle.orients(p5[x], 0xFF)) == 1
    en = default_masks(oracle, p5)
    en[x] = 0xFF
    assert len(oracle.solve(p5, [1] * 12, scott, enabled=en)) == 8 * 520
    rot = 0b01101001

    def rotations_only(shapes):
        out = []
        for s in shapes:
            seen, m = [], 0
            for t in range(8):
                if rot >> t & 1:
                    img = oracle.transform(s, t)
                    if img not in seen:
                        m |= 1 << t
                    seen.append(img)
            out.append(m)
        return out

    shapes4 = oracle.enum_free(4)
    for shapes, counts, rows in [(p5, [1] * 12, rect(5, 6, [(0, 0), (5, 4)])),
                                 (shapes4, [2] * 5, rect(4, 5)),
                                 (shapes4, [1, 2, 1, 2, 1], rect(5, 4))]:
        got = oracle.solve(shapes, counts, rows, enabled=rotations_only(shapes))
        assert len(got) == oracle.count(shapes, counts, rows, sym=rot)
    dom = [3, 1]  # domino with its placements off, monomino
    assert len(oracle.solve(dom, [8, 16], rect(4, 4), enabled=[0, 1])) == 1
    assert len(oracle.solve(dom, [8, 15], rect(4, 4), enabled=[0, 1])) == 0


def test_record_codec_round_trip(oracle):
    """parse_R<H>(assemble_R<H>(cfg)) restores cfg (the reference's
    bmark/record.cu property) for every height and id count it can hold."""
    rng = random.Random(5)
    cap = {7: 8, 6: 14, 5: 15, 4: 16, 3: 16, 2: 16, 1: 16}
    for H in range(1, 8):
        for _ in range(300):
            cnt = rng.randint(0, cap[H])
            ids = sorted(rng.sample(range(40), cnt))
            e = rng.getrandbits(8 * (H + 1))  # rows 0..H; row 0 just filled
            r, ea = oracle.assemble_R(H, e, cnt, ids)
            e2, c2, nm = oracle.parse_R(H, r, ea)
            assert e2 == e >> 8, H
            assert c2 == cnt
            assert nm[:cnt] == ids
            # positions the record holds beyond the ids read as padding
            assert all(v == 0xFF for v in nm[cnt:cap[H]])


def test_record_codec_layout(oracle):
    # record.h: R = {xaL, xaH, ex0, ex1, ex2}; ids fill ex0, ex1, ex2 then
    # the spare bytes of xaH (H 5, 6) or all of xaH (H <= 4)
    ids = list(range(1, 15))
    e = 0x77_66_55_44_33_22_11_00  # byte k = row k; row 0 just filled
    r, ea = oracle.assemble_R(6, e, 14, ids)
    assert ea == 0x11                                   # new current row
    assert r[0] == 0x55443322                           # new rows 1..4
    assert (r[1] >> 16) & 0xFF == 0x66                  # new row 5
    assert r[2] == 0x04030201 and r[3] == 0x08070605 and r[4] == 0x0C0B0A09
    assert r[1] & 0xFFFF == 0x0E0D and r[1] >> 24 == 14
    r, ea = oracle.assemble_R(7, e, 2, [3, 9])
    assert r[0] == 0x55443322 and r[1] == 0x7766 | (2 << 24)
    r, ea = oracle.assemble_R(4, 0x00_112233_44 << 8, 3, [7, 8, 9])
    assert r[0] == (0x112233 & 0xFFFFFF) | (3 << 24)
    assert r[2] == 0xFF090807 and r[1] == 0xFFFFFFFF


def test_step_matrix_maps_transform_back(oracle):
    """make_step's matrices (src/Piece.cpp:41-54, restated as
    or_step_matrix) against the transforms themselves (or_transform, pinned
    by the reference's compiled Shape::transforms): for transform t, the
    matrix M maps every cell (x, y) = (column, row) of the transformed shape
    back onto the original shape's cells, up to translation (input = M .
    output).  Shapes with no symmetry make a wrong sign, a swapped entry or
    a matrix listed under the wrong t fail, and the product's table
    (d4pieces) must equal the pinned one."""
    import ctypes

    def cells(s):
        return [(b % 8, b // 8) for b in range(64) if s >> b & 1]

    def norm(pts):
        mx = min(x for x, _ in pts)
        my = min(y for _, y in pts)
        return sorted((x - mx, y - my) for x, y in pts)

    # every free polyomino of 4-6 cells with no symmetry (8 distinct
    # transforms)
    asym = [s for n in (4, 5, 6) for s in oracle.enum_free(n)
            if len({oracle.transform(s, t) for t in range(8)}) == 8]
    assert len(asym) == 1 + 5 + 20  # L; F, L, N, P, Y; 20 hexominoes
    for s in asym:
        orig = norm(cells(s))
        for t in range(8):
            m = (ctypes.c_int * 4)()
            oracle.L.or_step_matrix(t, m)
            a, b, c, d = m
            assert sorted(map(abs, (a, b, c, d))) == [0, 0, 1, 1]
            back = [(a * x + b * y, c * x + d * y) for x, y in cells(oracle.transform(s, t))]
            assert norm(back) == orig, (hex(s), t, tuple(m))
            # no other listed matrix maps transform t back
            for u in range(8):
                if u == t:
                    continue
                mu = (ctypes.c_int * 4)()
                oracle.L.or_step_matrix(u, mu)
                au, bu, cu, du = mu
                wrong = [(au * x + bu * y, cu * x + du * y)
                         for x, y in cells(oracle.transform(s, t))]
                assert norm(wrong) != orig
    import d4pieces
    for t in range(8):
        m = (ctypes.c_int * 4)()
        oracle.L.or_step_matrix(t, m)
        assert tuple(d4pieces._STEP_MATRIX[t]) == tuple(m)


def _solve_l():
    import sys
    from conftest import ROOT
    p = os.path.join(ROOT, "oracle")
    if p not in sys.path:
        sys.path.insert(0, p)
    import solve_l
    return solve_l


def _to_frame(v8, L):
    """An 8x8 value (bit = 8r + c) in the L x L frame."""
    out = 0
    for b in range(64):
        if v8 >> b & 1:
            out |= 1 << (L * (b // 8) + b % 8)
    return out


def test_solve_l_matches_oracle(oracle):
    """The L x L restatement (oracle/solve_l.py) gives the C oracle's
    solutions, steps and order on boards up to 8 wide, in the 8x8 and the
    11x11 frame: random piece sets of 1-4 cells, multiplicities 1-2, symmetry
    masks D4 / C4 / D1_X / C1, and per-placement enabled masks that switch
    duplicates on and placements off."""
    S = _solve_l()
    rng = random.Random(17)
    pool = [1, 3, 7, 0x103] + [s for s in _free(oracle, 4)]
    checked = 0
    for _ in range(60):
        shapes = sorted(rng.sample(pool, rng.randint(1, 4)))
        counts = [rng.randint(1, 2) for _ in shapes]
        sym = rng.choice([0xFF, 0x69, 0x03, 0x01])
        nplace = bin(sym).count("1")
        enabled = [rng.getrandbits(nplace) | 1 for _ in shapes] if rng.random() < 0.5 else None
        w, h = rng.randint(2, 8), rng.randint(1, 4)
        rows = [rng.getrandbits(w) | (rng.getrandbits(w) if rng.random() < 0.7 else 0)
                for _ in range(h)]
        rows = [r | ((1 << w) - 1) * (rng.random() < 0.5) for r in rows]
        expect = oracle.solve(shapes, counts, rows, sym=sym, enabled=enabled)
        for L in (8, 11):
            lib = []
            for i, (s, c) in enumerate(zip(shapes, counts)):
                p = S.Piece(_to_frame(s, L), L, sym, c)
                if enabled is not None:
                    for q, pl in enumerate(p.placements):
                        pl[2] = bool(enabled[i] >> q & 1)
                lib.append(p)
            board = sum(_to_frame(r, L) << (L * i) for i, r in enumerate(rows))
            got = [[(pid, trs, y, x) for pid, trs, x, y, _ in sol]
                   for sol in S.solve(lib, board, L)]
            assert got == expect, (shapes, counts, sym, enabled, rows, L)
            first = S.solve(lib, board, L, single=True)
            assert [[(p_, t, y, x) for p_, t, x, y, _ in s] for s in first] == expect[:1]
        checked += len(expect)
    assert checked > 200


def _free(oracle, n):
    import d4pieces as d4
    return d4.free_polyominoes(n)


def test_solve_l_closed_forms_11_wide():
    """On 11-wide boards: dominoes on 2x11 tile F(12) = 144 ways, straight
    trominoes on 3x11 a(11) = 41 ways (a(n) = a(n-1) + a(n-3), a(0..2) = 1),
    and each tiling covers the board exactly once."""
    S = _solve_l()
    L = 11
    full = (1 << 11) - 1
    sols = S.solve([S.Piece(0b11, L, 0xFF, 11)], full | full << 11, L)
    assert len(sols) == 144
    for sol in sols:
        cover = 0
        for _, _, _, _, placed in sol:
            assert not cover & placed
            cover |= placed
        assert cover == full | full << 11
    a = [1, 1, 1]
    for n in range(3, 12):
        a.append(a[-1] + a[-3])
    sols = S.solve([S.Piece(0b111, L, 0xFF, 11)], full | full << 11 | full << 22, L)
    assert len(sols) == a[11] == 41


def test_solve_l_transposition():
    """A board and its transpose (11x11 frame) have equally many tilings
    under D4, and the transposed tilings are the transposes of the tilings."""
    S = _solve_l()
    L = 11
    rng = random.Random(23)
    shapes = [0b111, 0b11 | 0b1 << 11, 0b11 | 0b11 << 11]  # I3, L3, O4
    for _ in range(12):
        w, h = rng.randint(6, 11), rng.randint(2, 4)
        board = 0
        for r in range(h):
            for c in range(w):
                if rng.random() < 0.85:
                    board |= 1 << (r * L + c)
        tb = S.transform(board, 4, L, norm=False)
        lib = [S.Piece(s, L, 0xFF, 3) for s in shapes]
        a = {frozenset(p for *_, p in sol) for sol in S.solve(lib, board, L)}
        b = {frozenset(S.transform(p, 4, L, norm=False) for *_, p in sol)
             for sol in S.solve(lib, tb, L)}
        assert a == b
