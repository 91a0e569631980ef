"""The reference's embind surface (src/wasm.cpp:10-146, the `pieces` module
its web UI calls) over the device search of d4pieces: the same names, fields
and argument meanings, with solve8 answered by the GPU searcher instead of the
CPU backtracker.

    Shape8(value)          value, LEN, size, normalize(), canonical_form(forms),
                           classify, symmetry, transform0() .. transform7(),
                           top, left, width, height, bottom, right, test(r, c),
                           set(r, c), clear(r, c)            (src/Shape.hpp)
    Piece8(shape)          count, shape, placements[] of Placement8
                           {normal, enabled, duplicate}    (src/Piece.hpp:10-29)
    Step8 / Solution       piece_id, trs_id, a, b, c, d, x, y, shape / steps, map
    solve8(pieces, board, single) -> [Solution]           (src/Piece.cpp:65-102)
    shape_count(n), shape_at8(n, i)                       (src/known.hpp:9-18)
    SymmetryGroup, subgroup(a, b), groupProduct(a, b)     (src/Group.hpp:8-50)

solve11 / Shape11 (11x11 boards) are not offered: the device search takes
boards up to 8 columns wide.
"""
from enum import IntEnum
from typing import List, Sequence

import d4pieces as d4


class SymmetryGroup(IntEnum):
    """Named subgroups of D4 as transform masks (src/Group.hpp:8-19)."""
    C1 = 0b00000001
    C2 = 0b00001001
    C4 = 0b01101001
    D1_X = 0b00000011
    D1_Y = 0b00000101
    D1_P = 0b00010001
    D1_S = 0b10000001
    D2_XY = 0b00001111
    D2_PS = 0b10011001
    D4 = 0b11111111


def subgroup(a: int, b: int) -> bool:
    """a >= b: every transform of b is in a (src/Group.hpp:21-25)."""
    return (int(a) & int(b)) == int(b)


def _compose_table():
    """t_i after t_j as a transform index, read off an asymmetric shape."""
    probe = 0b1_00000111 | (1 << 17)  # rows: ###, #.., .#. -- no symmetry
    img = [d4.transform8(probe, t) for t in range(8)]
    return [[img.index(d4.transform8(img[j], i)) for j in range(8)] for i in range(8)]


_COMPOSE = _compose_table()


def groupProduct(a: int, b: int) -> int:  # noqa: N802 (the reference's name)
    """The smallest named group holding both: the closure of a | b under
    composition.  (The reference closes with a table of rules,
    src/Group.hpp:27-50, which leaves some products, e.g. C4 x D1_P, at a
    mask that is not a group; this returns the group, D4 there.)"""
    m = int(a) | int(b) | 1
    while True:
        n = m
        for i in range(8):
            for j in range(8):
                if m >> i & 1 and m >> j & 1:
                    n |= 1 << _COMPOSE[i][j]
        if n == m:
            return m
        m = n


class Shape8:
    """Shape<8> (src/Shape.hpp): an 8x8 bitboard, bit = row * 8 + col."""
    LEN = 8

    def __init__(self, value: int = 0):
        self.value = int(value) & ((1 << 64) - 1)

    def __eq__(self, other):
        return isinstance(other, Shape8) and self.value == other.value

    def __hash__(self):
        return hash(self.value)

    def __repr__(self):
        return f"Shape8({self.value:#x})"

    @property
    def size(self) -> int:
        return bin(self.value).count("1")

    def normalize(self) -> "Shape8":
        return Shape8(d4.normalize8(self.value))

    def _transform(self, t: int) -> "Shape8":
        return Shape8(d4.transform8(self.value, t))

    def transform0(self, norm: bool = True) -> "Shape8":
        return self._t(0, norm)

    def transform1(self, norm: bool = True) -> "Shape8":
        return self._t(1, norm)

    def transform2(self, norm: bool = True) -> "Shape8":
        return self._t(2, norm)

    def transform3(self, norm: bool = True) -> "Shape8":
        return self._t(3, norm)

    def transform4(self, norm: bool = True) -> "Shape8":
        return self._t(4, norm)

    def transform5(self, norm: bool = True) -> "Shape8":
        return self._t(5, norm)

    def transform6(self, norm: bool = True) -> "Shape8":
        return self._t(6, norm)

    def transform7(self, norm: bool = True) -> "Shape8":
        return self._t(7, norm)

    def _t(self, t: int, norm: bool) -> "Shape8":
        """Transform t; without norm the image stays where the 8x8 frame's
        flips put it (Shape::transform<..>(false))."""
        if norm:
            return self._transform(t)
        sw, fx, fy = t & 4, t & 1, t & 2
        out = 0
        for b in range(64):
            if self.value >> b & 1:
                y, x = divmod(b, 8)
                yy, xx = (7 - y if fy else y), (7 - x if fx else x)
                r, c = (xx, yy) if sw else (yy, xx)
                out |= 1 << (r * 8 + c)
        return Shape8(out)

    def canonical_form(self, forms: int = 0xFF) -> "Shape8":
        """Minimum over the transforms in `forms` of the normalised images
        (all ones when `forms` is empty)."""
        return Shape8(min([d4.transform8(self.value, t) for t in range(8) if forms >> t & 1],
                          default=(1 << 64) - 1))

    @property
    def symmetry(self) -> int:
        """Mask of the transforms that map the normalised shape onto itself."""
        n = d4.normalize8(self.value)
        return sum(1 << t for t in range(8) if d4.transform8(self.value, t) == n)

    @property
    def classify(self) -> SymmetryGroup:
        return SymmetryGroup(self.symmetry)

    # Margins and extents counted as the reference counts them
    # (src/Shape.hpp:126-168): shift the whole 64-bit value until its first
    # column (row) is occupied, then count the shifts while it stays so.  A
    # shape spanning all 8 columns therefore counts on into the cells the
    # shift brings over from the next row, and right() = 8 - left - width
    # wraps as an unsigned 64-bit value; results are kept identical.
    _COL0, _ROW0, _M64 = 0x0101010101010101, 0xFF, (1 << 64) - 1

    def _run(self, first: int, step: int):
        v, skip, run = self.value, 0, 0
        if not v:
            return 8, 0
        while not v & first:
            v >>= step
            skip += 1
        while v & first:
            v >>= step
            run += 1
        return skip, run

    @property
    def left(self) -> int:
        return self._run(self._COL0, 1)[0]

    @property
    def width(self) -> int:
        return self._run(self._COL0, 1)[1]

    @property
    def right(self) -> int:
        skip, run = self._run(self._COL0, 1)
        return (8 - skip - run) & self._M64

    @property
    def top(self) -> int:
        return self._run(self._ROW0, 8)[0]

    @property
    def height(self) -> int:
        return self._run(self._ROW0, 8)[1]

    @property
    def bottom(self) -> int:
        skip, run = self._run(self._ROW0, 8)
        return (8 - skip - run) & self._M64

    def test(self, row: int, col: int) -> bool:
        return bool(self.value >> (row * 8 + col) & 1)

    def set(self, row: int, col: int) -> "Shape8":
        return Shape8(self.value | 1 << (row * 8 + col))

    def clear(self, row: int, col: int) -> "Shape8":
        return Shape8(self.value & ~(1 << (row * 8 + col)))


class Placement8:
    """Piece<8>::Placement: a transform's normalised shape, and whether the
    search uses it (off for duplicates of an earlier transform)."""
    __slots__ = ("normal", "enabled", "duplicate")

    def __init__(self, normal: Shape8, enabled: bool, duplicate: bool):
        self.normal, self.enabled, self.duplicate = normal, enabled, duplicate


class Piece8:
    """Piece<8>(shape) (src/Piece.cpp:12-22): count 1, and one placement per
    D4 transform in order, duplicates disabled."""

    def __init__(self, shape, sym: int = 0xFF):
        self.shape = shape if isinstance(shape, Shape8) else Shape8(shape)
        self.count = 1
        self.placements: List[Placement8] = []
        for t in range(8):
            if sym >> t & 1:
                n = self.shape._transform(t)
                dup = any(p.normal == n for p in self.placements)
                self.placements.append(Placement8(n, not dup, dup))
        self.sym = sym


Step8 = d4.FullStep


class Solution(d4.Solution):
    """Solution<8>: steps and the 8x8 map (step index per cell, -1 free)."""


def solve8(pieces: Sequence[Piece8], board: Shape8, single: bool = False) -> List[Solution]:
    """solve(lib, board, single) (src/Piece.cpp:65-102): every tiling of the
    board's set cells by the pieces (each at most `count` times, enabled
    placements only), in the CPU searcher's order, or its first one; solved
    on the device (enum mode).

    Enabled duplicates (placements whose shape equals an earlier enabled
    one) are searched once and expanded afterwards: a duplicate leads to the
    same cells and the same subtree, and the CPU order is lexicographic in
    (piece id, placement index) step by step, since at the anchor (the first
    open cell) each placement fits in at most one translation."""
    pieces = list(pieces)
    if not pieces:
        raise ValueError("solve8 needs at least one piece")
    sym = pieces[0].sym
    if any(p.sym != sym for p in pieces):
        raise ValueError("all pieces must share one symmetry mask")
    shapes = [d4.normalize8(p.shape.value) for p in pieces]
    masks, dups = [], []
    for p in pieces:
        rep, m = {}, 0  # normal shape -> enabled placement indices, in order
        for t, pl in enumerate(p.placements):
            if pl.enabled:
                if pl.normal.value not in rep:
                    rep[pl.normal.value] = []
                    m |= 1 << t
                rep[pl.normal.value].append(t)
        masks.append(m)
        dups.append({ts[0]: ts for ts in rep.values()})
    bv = board.value if isinstance(board, Shape8) else int(board)
    rows = [bv >> (8 * r) & 0xFF for r in range(8)]
    while rows and not rows[-1]:
        rows.pop()
    if not rows:
        # the empty board has one tiling, with no steps
        return [Solution([], 8)]
    lib = d4.PieceLibrary(shapes, [p.count for p in pieces], sym, 8, enabled=masks)
    s = d4.CudaSearcher(lib, rows=rows, mode=d4.MODE_ENUM)
    s.run()
    raws = s.solutions(single=single)
    if any(len(v) > 1 for d in dups for v in d.values()) and not single:
        import itertools
        out = []
        for raw in raws:
            choices = [dups[pid][trs] for pid, trs, _, _ in raw]
            for pick in itertools.product(*choices):
                out.append([(pid, t, y, x) for (pid, _, y, x), t in zip(raw, pick)])
        out.sort(key=lambda r: [(pid, t) for pid, t, _, _ in r])
        raws = out
    return [Solution(d4.expand_solution(shapes, raw, 8, sym).steps, 8) for raw in raws]


def shape_count(n: int) -> int:
    """Number of free polyominoes with n cells (the size of known_shapes[n])."""
    return len(d4.free_polyominoes(n)) if 1 <= n <= 8 else 0


def shape_at8(n: int, i: int) -> Shape8:
    """known_shapes[n][i] (src/known.hpp:15-18)."""
    return Shape8(d4.free_polyominoes(n)[i])
