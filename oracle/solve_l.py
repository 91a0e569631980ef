"""TEST INFRASTRUCTURE ONLY -- a pure-Python restatement of the reference's
CPU backtracker on an L x L frame, for small boards (tests/ use it as the
checker of the 11x11 solve; nothing in the product imports it).

Follows, for any frame size L (bit = row * L + col):
  Shape<L>::transform / transforms / normalize   src/Shape.cpp:6-41, src/Shape.hpp:182-191
  Shape<L>::bottom / right                       src/Shape.hpp:126-168
  Piece<L>::Piece(shape, sym)                    src/Piece.cpp:12-22
  Piece<L>::cover(pos, func)                     src/Piece.inl:5-24
  make_step, Solution<L>                         src/Piece.cpp:24-54
  solve(lib, board, single)                      src/Piece.cpp:65-102

Pinned (tests/test_oracle.py): against the C oracle's solve (itself pinned by
closed forms and brute force) on 8-wide boards in both frames, by the domino
Fibonacci and straight-tromino counts on 11-wide boards, and by transposition
invariance (a board and its transpose tile alike under D4).
"""
from typing import List, Sequence, Tuple

STEP_MATRIX = [(1, 0, 0, 1), (-1, 0, 0, 1), (1, 0, 0, -1), (-1, 0, 0, -1),
               (0, 1, 1, 0), (0, -1, 1, 0), (0, 1, -1, 0), (0, -1, -1, 0)]


def cells(v: int, L: int):
    while v:
        b = (v & -v).bit_length() - 1
        yield divmod(b, L)
        v &= v - 1


def normalize(v: int, L: int) -> int:
    if not v:
        return 0
    row0, col0 = (1 << L) - 1, sum(1 << (L * i) for i in range(L))
    while not v & row0:
        v >>= L
    while not v & col0:
        v >>= 1
    return v


def transform(v: int, t: int, L: int, norm: bool = True) -> int:
    """transform<Swap = t & 4, FlipX = t & 1, FlipY = t & 2>(norm)."""
    if not v:
        return 0
    sw, fx, fy = t & 4, t & 1, t & 2
    out = 0
    for r in range(L):
        for c in range(L):
            ir, ic = (c, r) if sw else (r, c)
            if fx:
                ic = L - ic - 1
            if fy:
                ir = L - ir - 1
            if v >> (ir * L + ic) & 1:
                out |= 1 << (r * L + c)
    return normalize(out, L) if norm else out


def extent(v: int, L: int) -> Tuple[int, int]:
    """(bottom, right) margins of a normalised shape (rows / columns free
    below and to the right of it in the frame)."""
    rows = [r for r, _ in cells(v, L)]
    cols = [c for _, c in cells(v, L)]
    return L - (max(rows) + 1), L - (max(cols) + 1)


class Piece:
    """Piece<L>(shape, sym): count 1, placements (normal, max, enabled,
    duplicate) for the transforms in sym, in order."""

    def __init__(self, shape: int, L: int, sym: int = 0xFF, count: int = 1):
        self.L, self.count, self.canonical = L, count, shape
        self.placements = []
        for t in range(8):
            if sym >> t & 1:
                n = transform(shape, t, L)
                dup = any(p[0] == n for p in self.placements)
                self.placements.append([n, extent(n, L), not dup, dup])

    def cover(self, pos, func) -> bool:
        ty, tx = pos
        for trs, (n, (my, mx), enabled, _) in enumerate(self.placements):
            if not enabled:
                continue
            for by, bx in cells(n, self.L):
                if bx > tx or bx + mx < tx or by > ty or by + my < ty:
                    continue
                x, y = tx - bx, ty - by
                if func(n << (self.L * y + x), trs, (y, x)):
                    return True
        return False


def solve(lib: Sequence[Piece], board: int, L: int, single: bool = False
          ) -> List[List[Tuple[int, int, int, int, int]]]:
    """Every tiling (or the first) as a list of steps (piece_id, trs_id, x,
    y, placed shape) in the searcher's order."""
    sols, hist = [], []
    used = [0] * len(lib)
    size = [bin(p.canonical).count("1") for p in lib]
    max_tiles = [sum(s * p.count for s, p in zip(size, lib))]

    def min_tiles():
        m = None
        for p, u, s in zip(lib, used, size):
            if u < p.count:
                m = s if m is None else min(m, s)
        return m

    def f(open_tiles: int) -> bool:
        n = bin(open_tiles).count("1")
        if n > max_tiles[0]:
            return False
        if not open_tiles:
            sols.append(list(hist))
            return single
        mt = min_tiles()
        if mt is None or n < mt:
            return False
        b = (open_tiles & -open_tiles).bit_length() - 1
        pos = divmod(b, L)
        for pid, p in enumerate(lib):
            if used[pid] == p.count:
                continue
            used[pid] += 1
            max_tiles[0] -= size[pid]

            def take(placed, trs, tra, pid=pid):
                if placed & ~open_tiles:
                    return False
                hist.append((pid, trs, tra[1], tra[0], placed))
                r = f(open_tiles & ~placed)
                hist.pop()
                return r

            if p.cover(pos, take):
                return True
            max_tiles[0] += size[pid]
            used[pid] -= 1
        return False

    f(board)
    return sols
