"""Seeded synthetic workloads for the tiling search.

Every random choice comes from std::mt19937_64 (the generator the reference's
benches use, bmark/sn.cu:54, bmark/push_nm.cu:229), restated here so the
instances are reproducible from the seed alone, with a portable bounded-int
mapping (lo + floor(x * n / 2^64)) instead of std::uniform_int_distribution,
whose output differs between standard libraries.

Workloads (boards are lists of 8-bit row masks, bit c = column c open):
  plain_6x10()        12 free pentominoes x1 on 6 wide x 10 rows
  seeded_8x8(n, seed) 8x8 with 4 blocked cells, every open region's area a
                      multiple of 5 (first n accepted draws)
  large_8x15()        12 free pentominoes x2 on 8 wide x 15 rows
  uneven_instances()  4096 puzzles, each with its own 6-14 polyominoes of
                      sizes 3-6 (multiplicity 1-2) and a 4-8 wide board
  calendar_boards()   the reference's calendar puzzle: 12 x 31 boards, 8 pieces
  board_configs(text) the reference's board-file region format, expanded
  request_stream(...) Poisson arrivals of 8x8 boards with 30-60 open cells in
                      one region (area a multiple of 5), 70% first-solution
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

MASK64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (Matsumoto-Nishimura 64-bit Mersenne Twister)."""

    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed: int = 5489):
        mt = [0] * self.N
        mt[0] = seed & MASK64
        for i in range(1, self.N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & MASK64
        self.mt, self.idx = mt, self.N

    def _twist(self) -> None:
        mt, N, M = self.mt, self.N, self.M
        for i in range(N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[i] = mt[(i + M) % N] ^ xa
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK64

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + ((self() * (hi - lo + 1)) >> 64)

    def uniform(self) -> float:
        return (self() >> 11) * (1.0 / (1 << 53))


def pentominoes() -> List[int]:
    from d4pieces import free_polyominoes
    return free_polyominoes(5)


def regions(rows: Sequence[int], width: int = 8) -> List[int]:
    """Areas of the 4-connected open regions of a board."""
    cells = {(r, c) for r, v in enumerate(rows) for c in range(width) if (v >> c) & 1}
    areas = []
    while cells:
        stack = [cells.pop()]
        a = 0
        while stack:
            r, c = stack.pop()
            a += 1
            for q in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
                if q in cells:
                    cells.remove(q)
                    stack.append(q)
        areas.append(a)
    return areas


def plain_6x10() -> Tuple[List[int], List[int], List[int], int]:
    """(shapes, counts, rows, width) of 12 pentominoes x1 on 6x10."""
    return pentominoes(), [1] * 12, [0x3F] * 10, 6


def large_8x15() -> Tuple[List[int], List[int], List[int], int]:
    """(shapes, counts, rows, width) of 12 pentominoes x2 on 8x15."""
    return pentominoes(), [2] * 12, [0xFF] * 15, 8


def _blocked_board(rng: MT19937_64, n_blocked: int) -> List[int]:
    cells: List[int] = []
    while len(cells) < n_blocked:
        v = rng.randint(0, 63)
        if v not in cells:
            cells.append(v)
    rows = [0xFF] * 8
    for v in cells:
        rows[v // 8] &= 0xFF ^ (1 << (v % 8))
    return rows


def seeded_8x8(n: int = 256, seed: int = 2) -> List[List[int]]:
    """First n 8x8 boards with 4 distinct uniform blocked cells whose open
    regions all have area divisible by 5."""
    rng = MT19937_64(seed)
    out: List[List[int]] = []
    while len(out) < n:
        rows = _blocked_board(rng, 4)
        if all(a % 5 == 0 for a in regions(rows)):
            out.append(rows)
    return out


def centre_hole_8x8() -> List[int]:
    """8x8 minus the central 2x2 (Scott 1958: 520 pentomino tilings, 65 x 8)."""
    return [0xFF, 0xFF, 0xFF, 0xE7, 0xE7, 0xFF, 0xFF, 0xFF]


def grown_region_8x8(rng: MT19937_64, free: int) -> List[int]:
    """An 8x8 board whose open cells are one 4-connected region of `free`
    cells, grown from a uniform seed cell by adding uniform frontier cells."""
    start = rng.randint(0, 63)
    cells = {start}
    frontier = []

    def nbrs(v):
        r, c = divmod(v, 8)
        if r > 0:
            yield v - 8
        if r < 7:
            yield v + 8
        if c > 0:
            yield v - 1
        if c < 7:
            yield v + 1

    frontier.extend(x for x in nbrs(start))
    while len(cells) < free:
        v = frontier[rng.randint(0, len(frontier) - 1)]
        if v in cells:
            frontier.remove(v)
            continue
        cells.add(v)
        frontier.extend(x for x in nbrs(v) if x not in cells)
    rows = [0] * 8
    for v in cells:
        rows[v // 8] |= 1 << (v % 8)
    return rows


def uneven_instances(n: int = 4096, seed: int = 3):
    """n puzzles of uneven size: list of (shapes, counts, rows, width).

    Piece set: k ~ U{6..14} distinct shapes drawn uniformly (rejecting
    repeats) from the 54 free polyominoes of sizes 3..6 (in size, then
    canonical-value order), kept in that order; multiplicity U{1, 2} each.
    Board: width W ~ U{4..8}, H = ceil(A / W) rows for the total piece area
    A; the W*H - A (< W) surplus cells, drawn uniformly and distinct over the
    row-major cell index r*W + c, are blocked, so the open area equals A and a
    tiling uses every piece its full multiplicity."""
    from d4pieces import free_polyominoes
    pool = [s for k in (3, 4, 5, 6) for s in free_polyominoes(k)]
    rng = MT19937_64(seed)
    out = []
    for _ in range(n):
        k = rng.randint(6, 14)
        idx: List[int] = []
        while len(idx) < k:
            v = rng.randint(0, len(pool) - 1)
            if v not in idx:
                idx.append(v)
        idx.sort()
        shapes = [pool[i] for i in idx]
        counts = [rng.randint(1, 2) for _ in shapes]
        area = sum(c * bin(s).count("1") for s, c in zip(shapes, counts))
        W = rng.randint(4, 8)
        H = -(-area // W)
        blocked: List[int] = []
        while len(blocked) < W * H - area:
            v = rng.randint(0, W * H - 1)
            if v not in blocked:
                blocked.append(v)
        rows = [(1 << W) - 1] * H
        for v in blocked:
            rows[v // W] &= ~(1 << (v % W)) & 0xFF
        out.append((shapes, counts, rows, W))
    return out


from d4pieces import normalize8, transform8  # noqa: E402  (geometry helpers)


# The reference's calendar puzzle (src/main.cpp:14-42): 8 pieces, a board of
# 6 + 6 month cells and 31 day cells (the 8x8 value below turned by 180
# degrees), and for each month/day the board with those two cells removed.
CALENDAR_PIECES = [
    0b11110000_10000000,           # L
    0b11100000_00110000,           # Z
    0b01100000_01000000_11000000,  # S
    0b10000000_10000000_11100000,  # L_
    0b11100000_11100000,           # rect
    0b11100000_11000000,           # rect-
    0b11100000_10100000,           # C
    0b11110000_00100000,           # |-
]
CALENDAR_BOARD = 0b11111100_11111100_11111110_11111110_11111110_11111110_11100000


def calendar_boards():
    """(shapes, counts, [(month, day, rows)], width): the 12 x 31 calendar
    boards, rows as 8-bit masks (width 7)."""
    shapes = [normalize8(s) for s in CALENDAR_PIECES]
    board = transform8(CALENDAR_BOARD, 3)  # flip X and Y (src/main.cpp:37)
    out = []
    for m in range(1, 13):
        for d in range(1, 32):
            b = board & ~(1 << ((m - 1) // 6 * 8 + (m - 1) % 6))
            b &= ~(1 << (((d - 1) // 7 + 2) * 8 + (d - 1) % 7))
            rows = [(b >> (8 * r)) & 0xFF for r in range(8)]
            while rows and not rows[-1]:
                rows.pop()
            out.append((m, d, rows))
    return shapes, [1] * len(shapes), out, 7


def board_configs(text: str, width: int = 8):
    """The board-file region format (src/board.hpp:15-57): every letter or
    digit marks a cell of the region of that character, other characters
    are read as Shape text ('#', '*', '1', '@', 'X', '%' filled; ' ', '.',
    '0', '-', '_' empty; newline = next row).  The base board fills every
    region cell; one configuration removes one cell of every region (in
    region order of first appearance, cells in bit order).  Returns
    (count, list of row-mask lists)."""
    def valid(ch):
        return ch.isascii() and ch.isalnum()

    def parse(sv):
        if sv.startswith("\n"):
            sv = sv[1:]
        value, row, bit = 0, 0, 0
        for ch in sv:
            if ch in " 0.-_":
                bit += 1
            elif ch in "*1@#X%":
                value |= 1 << (row * width + bit)
                bit += 1
            elif ch == "\n":
                row, bit = row + 1, 0
        return value

    keys = []
    for ch in text:
        if valid(ch) and ch not in keys:
            keys.append(ch)
    base = parse("".join("#" if valid(ch) else ch for ch in text))
    regions = [parse("".join(("#" if ch == k else ".") if valid(ch) else ch for ch in text))
               for k in keys]
    configs = []

    def rec(i, cur):
        if i == len(regions):
            rows = [(cur >> (width * r)) & ((1 << width) - 1) for r in range(64 // width)]
            while rows and not rows[-1]:
                rows.pop()
            configs.append(rows)
            return
        for b in range(64):
            if regions[i] >> b & 1:
                rec(i + 1, cur & ~(1 << b))

    rec(0, base)
    count = 1
    for r in regions:
        count *= bin(r).count("1")
    return count, configs


def request_stream(n: int, rate: float, seed: int = 4):
    """n requests with Poisson(rate) arrivals: list of (t_arrival_s, rows,
    single).  Each board is 8x8 with F ~ U{30..64} open cells rounded down to
    a multiple of 5 (so every open region's area is a multiple of 5), the
    open cells forming one connected region grown from a uniform seed cell.
    (Uniform blocked cells with a region-area filter almost never pass when
    20 or more cells are blocked.)  70% of the requests ask for the first
    solution only."""
    import math
    rng = MT19937_64(seed)
    t = 0.0
    out = []
    for _ in range(n):
        u = rng.uniform()
        t += -math.log(1.0 - u) / rate
        free = rng.randint(30, 64) // 5 * 5
        rows = grown_region_8x8(rng, free)
        single = rng.uniform() < 0.7
        out.append((t, rows, single))
    return out
