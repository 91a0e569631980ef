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

    Shape11, Piece11, Placement11, Step11, Solution11, solve11, shape_at11
                           the same on the 11x11 frame (121-bit values,
                           bit = row * 11 + col); solve11 runs on the device
                           when the board's cells span at most 8 columns
                           (searched as they lie) or at most 8 rows (searched
                           transposed, solutions mapped back and put in the
                           CPU searcher's order); a board spanning more than 8
                           of both is refused (ValueError).
"""
import itertools
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


class Shape11:
    """Shape<11> (src/Shape.hpp): an 11x11 bitboard, bit = row * 11 + col,
    in a 128-bit word masked to 121 bits.  Transforms, margins and extents as
    the reference computes them for LEN = 11 (src/Shape.cpp:6-79,
    src/Shape.hpp:126-191)."""
    LEN = 11
    _FULL = (1 << 121) - 1
    _ROW0 = (1 << 11) - 1
    _COL0 = sum(1 << (11 * i) for i in range(11))
    _M64 = (1 << 64) - 1

    def __init__(self, value: int = 0):
        self.value = int(value) & self._FULL

    def __eq__(self, other):
        return isinstance(other, Shape11) and self.value == other.value

    def __hash__(self):
        return hash(self.value)

    def __repr__(self):
        return f"Shape11({self.value:#x})"

    @property
    def size(self) -> int:
        return bin(self.value).count("1")

    def normalize(self) -> "Shape11":
        v = self.value
        if not v:
            return Shape11(0)
        while not v & self._ROW0:
            v >>= 11
        while not v & self._COL0:
            v >>= 1
        return Shape11(v)

    def _t(self, t: int, norm: bool) -> "Shape11":
        """transform<Swap = t & 4, FlipX = t & 1, FlipY = t & 2>(norm): out
        cell (r, c) reads (c, r) when swapped, then the flips mirror the
        column and the row within the 11x11 frame."""
        v = self.value
        if not v:
            return Shape11(0)
        sw, fx, fy = t & 4, t & 1, t & 2
        out = 0
        while v:
            b = (v & -v).bit_length() - 1
            v &= v - 1
            y, x = divmod(b, 11)
            yy, xx = (10 - y if fy else y), (10 - x if fx else x)
            r, c = (xx, yy) if sw else (yy, xx)
            out |= 1 << (r * 11 + c)
        sh = Shape11(out)
        return sh.normalize() if norm else sh

    def transform0(self, norm: bool = True) -> "Shape11":
        return self._t(0, norm)

    def transform1(self, norm: bool = True) -> "Shape11":
        return self._t(1, norm)

    def transform2(self, norm: bool = True) -> "Shape11":
        return self._t(2, norm)

    def transform3(self, norm: bool = True) -> "Shape11":
        return self._t(3, norm)

    def transform4(self, norm: bool = True) -> "Shape11":
        return self._t(4, norm)

    def transform5(self, norm: bool = True) -> "Shape11":
        return self._t(5, norm)

    def transform6(self, norm: bool = True) -> "Shape11":
        return self._t(6, norm)

    def transform7(self, norm: bool = True) -> "Shape11":
        return self._t(7, norm)

    def transforms(self, norm: bool = True) -> List["Shape11"]:
        return [self._t(t, norm) for t in range(8)]

    def canonical_form(self, forms: int = 0xFF) -> "Shape11":
        """Minimum over the transforms in `forms` of the normalised images
        (all 121 bits set when `forms` is empty)."""
        return Shape11(min([self._t(t, True).value for t in range(8) if forms >> t & 1],
                           default=self._FULL))

    @property
    def symmetry(self) -> int:
        n = self.normalize()
        return sum(1 << t for t in range(8) if self._t(t, True) == n)

    @property
    def classify(self) -> SymmetryGroup:
        return SymmetryGroup(self.symmetry)

    def _run(self, first: int, step: int):
        # (as Shape8._run: shift the whole word, skip, then count the run)
        v, skip, run = self.value, 0, 0
        if not v:
            return 11, 0
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
        return (11 - skip - run) & self._M64

    @property
    def top(self) -> int:
        return self._run(self._ROW0, 11)[0]

    @property
    def height(self) -> int:
        return self._run(self._ROW0, 11)[1]

    @property
    def bottom(self) -> int:
        skip, run = self._run(self._ROW0, 11)
        return (11 - skip - run) & self._M64

    def test(self, row: int, col: int) -> bool:
        return bool(self.value >> (row * 11 + col) & 1)

    def set(self, row: int, col: int) -> "Shape11":
        return Shape11(self.value | 1 << (row * 11 + col))

    def clear(self, row: int, col: int) -> "Shape11":
        return Shape11(self.value & ~(1 << (row * 11 + col)))

    def cells(self):
        v = self.value
        while v:
            b = (v & -v).bit_length() - 1
            v &= v - 1
            yield divmod(b, 11)


class Placement11:
    """Piece<11>::Placement: normal shape, enabled, duplicate."""
    __slots__ = ("normal", "enabled", "duplicate")

    def __init__(self, normal: Shape11, enabled: bool, duplicate: bool):
        self.normal, self.enabled, self.duplicate = normal, enabled, duplicate


class Piece11:
    """Piece<11>(shape) (src/Piece.cpp:12-22): count 1, one placement per
    transform in `sym`, duplicates disabled."""

    def __init__(self, shape, sym: int = 0xFF):
        self.shape = shape if isinstance(shape, Shape11) else Shape11(shape)
        self.count = 1
        self.placements: List[Placement11] = []
        for t in range(8):
            if sym >> t & 1:
                n = self.shape._t(t, True)
                dup = any(p.normal == n for p in self.placements)
                self.placements.append(Placement11(n, not dup, dup))
        self.sym = sym


class Step11:
    """Step<11>: piece_id, trs_id, the matrix a, b, c, d that make_step picks
    by the placement index, the translation x, y and the placed Shape11."""
    __slots__ = ("piece_id", "trs_id", "a", "b", "c", "d", "x", "y", "shape")

    def __init__(self, piece_id, trs_id, x, y, shape: Shape11):
        self.piece_id, self.trs_id, self.x, self.y, self.shape = piece_id, trs_id, x, y, shape
        self.a, self.b, self.c, self.d = d4._STEP_MATRIX[trs_id]

    def __repr__(self):
        return f"Step11(piece={self.piece_id}, trs={self.trs_id}, x={self.x}, y={self.y})"


class Solution11:
    """Solution<11>: steps, and map[row][col] (11 x 11) = index of the step
    covering the cell, -1 for none (src/Piece.cpp:24-39)."""

    def __init__(self, steps: List[Step11]):
        self.steps = steps
        self.map = [[next((i for i, st in enumerate(steps) if st.shape.test(r, c)), -1)
                     for c in range(11)] for r in range(11)]


def _to8(v11: int) -> int:
    """An 11-frame shape whose cells lie in rows and columns 0..7, as an 8x8 value."""
    out = 0
    for r, c in Shape11(v11).cells():
        if r > 7 or c > 7:
            raise ValueError("solve11: a piece larger than 8x8")
        out |= 1 << (8 * r + c)
    return out


def _transpose_idx(t: int) -> int:
    """j with transpose(T_t(s)) == T_j(transpose(s)) for every shape s."""
    return _COMPOSE[4][_COMPOSE[t][4]]


def solve11(pieces: Sequence[Piece11], board: Shape11, single: bool = False
            ) -> List[Solution11]:
    """solve(lib, board, single) (src/Piece.cpp:65-102) on the 11x11 frame,
    on the device.  A board whose cells span at most 8 columns is searched
    as it lies (shifted left), which keeps the CPU searcher's order; one
    spanning more than 8 columns but at most 8 rows is searched transposed
    (pieces transposed, each placement index mapped through the transpose),
    its solutions mapped back and sorted into the CPU searcher's order (the
    steps of a solution re-ordered by the lowest open cell of the board as
    it lies; solutions in lexicographic order of their (piece id, placement
    index) sequences, which is the depth-first order).  single=True returns
    the first solution in that order."""
    pieces = list(pieces)
    if not pieces:
        raise ValueError("solve11 needs at least one piece")
    sym = pieces[0].sym
    if any(p.sym != sym for p in pieces):
        raise ValueError("all pieces must share one symmetry mask")
    bv = board.value if isinstance(board, Shape11) else Shape11(board).value
    if not bv:
        return [Solution11([])]
    cl = list(Shape11(bv).cells())
    r_lo, r_hi = min(r for r, _ in cl), max(r for r, _ in cl)
    c_lo, c_hi = min(c for _, c in cl), max(c for _, c in cl)
    if c_hi - c_lo < 8:
        transposed = False
    elif r_hi - r_lo < 8:
        transposed = True
    else:
        raise ValueError("solve11: the board's cells span more than 8 rows and more than 8 "
                         "columns; the device search takes boards up to 8 columns wide "
                         "(transposed when the rows are)")
    order = [t for t in range(8) if sym >> t & 1]     # placement index -> transform
    if transposed:
        tsym = sum(1 << _transpose_idx(t) for t in order)
        torder = [t for t in range(8) if tsym >> t & 1]
        # placement index on the transposed board -> placement index as given
        back = [order.index(next(t for t in order if _transpose_idx(t) == j)) for j in torder]
    else:
        tsym, back = sym, list(range(len(order)))
    fwd = {q: i for i, q in enumerate(back)}
    shapes, masks, dups = [], [], []
    for p in pieces:
        canon = p.shape.normalize()
        shapes.append(_to8((canon._t(4, True) if transposed else canon).value))
        rep, m = {}, 0  # normal shape -> enabled placement indices (as given), in order
        for q, pl in enumerate(p.placements):
            if pl.enabled:
                if pl.normal.value not in rep:
                    rep[pl.normal.value] = []
                    m |= 1 << fwd[q]
                rep[pl.normal.value].append(q)
        masks.append(m)
        dups.append({fwd[qs[0]]: qs for qs in rep.values()})
    if transposed:
        rows = [sum(1 << (r - r_lo) for r in range(r_lo, r_hi + 1) if bv >> (r * 11 + c) & 1)
                for c in range(c_lo, c_hi + 1)]
    else:
        rows = [(bv >> (r * 11)) >> c_lo & 0xFF for r in range(r_hi + 1)]
    lib = d4.PieceLibrary(shapes, [p.count for p in pieces], tsym, 8, enabled=masks)
    s = d4.CudaSearcher(lib, rows=rows, mode=d4.MODE_ENUM)
    s.run()
    raws = s.solutions(single=single and not transposed)
    sols = []
    for raw in raws:
        placed = []  # (piece id, placement indices as given, y, x) on the board as it lies
        for pid, trs, y, x in raw:
            qs = dups[pid][trs]
            if transposed:
                placed.append((pid, qs, x + r_lo, y + c_lo))
            else:
                placed.append((pid, qs, y, x + c_lo))
        for pick in itertools.product(*[qs for _, qs, _, _ in placed]):
            steps = []
            for (pid, _, y, x), q in zip(placed, pick):
                shape = Shape11(pieces[pid].placements[q].normal.value << (11 * y + x))
                steps.append(Step11(pid, q, x, y, shape))
            sols.append(steps)
    if transposed or any(len(v) > 1 for d in dups for v in d.values()):
        sols = [_dfs_order(steps, bv) for steps in sols]
        sols.sort(key=lambda st: [(x.piece_id, x.trs_id) for x in st])
        if single:
            sols = sols[:1]
    return [Solution11(st) for st in sols]


def _dfs_order(steps: List[Step11], board: int) -> List[Step11]:
    """The steps in the order the CPU searcher places them: each covers the
    lowest open cell left by the ones before."""
    out, open_ = [], board
    left = list(steps)
    while open_:
        b = (open_ & -open_).bit_length() - 1
        i = next(i for i, st in enumerate(left) if st.shape.value >> b & 1)
        st = left.pop(i)
        out.append(st)
        open_ &= ~st.shape.value
    return out


def shape_at11(n: int, i: int) -> Shape11:
    """shape_at<11>(n, i) (src/known.hpp:15-18): known_shapes[n][i] in the
    11x11 frame."""
    v8 = d4.free_polyominoes(n)[i]
    out = 0
    for b in range(64):
        if v8 >> b & 1:
            r, c = divmod(b, 8)
            out |= 1 << (11 * r + c)
    return Shape11(out)


def shape_count(n: int) -> int:
    """Number of free polyominoes with n cells (the size of known_shapes[n])."""
    return len(d4.free_polyominoes(n)) if 1 <= n <= 8 else 0


def shape_at8(n: int, i: int) -> Shape8:
    """known_shapes[n][i] (src/known.hpp:15-18)."""
    return Shape8(d4.free_polyominoes(n)[i])
