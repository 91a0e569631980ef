"""Oracle for piece-set names -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product (libd4.so and the host
mirror) never does.  It restates, by brute force, the order in which the
reference names piece sets (src/naming.hpp:8-14, src/naming.cpp:35-81,
src/naming.inl:6-72):

* the pieces are grouped by size m = min_m .. max_m; piece id = position in
  that class-by-class list (Naming::name_piece, naming.hpp:20-24);
* a named set takes at most one of each piece, between min_n and max_n
  pieces, covering tgt cells;
* sets are ordered first by their partition -- the vector of pieces taken
  per size class -- in lexicographic order (the constructor's depth-first
  enumeration, naming.cpp:42-61, tries 0, 1, 2 ... pieces of the smallest
  size first, and name() finds the partition by binary search,
  naming.inl:15, which needs that order);
* within a partition, by the chosen indices of the smallest size class,
  then the next, each compared as the sorted index tuple in lexicographic
  order (name_binomial, naming.inl:31-44: a set that takes index i comes
  before every set with the same earlier choices that skips it).

Here every subset of the pieces is listed, filtered and sorted by that key;
the product computes ranks directly.  Small cases only (2^pieces subsets).

Pins (tests/test_naming.py): the reference's own test (naming.cpp:83-117:
every name resolves to a set of 20 cells that names back to itself); the
set count against a generating-function count; one size class against
itertools.combinations (lexicographic order by definition).
"""
from itertools import combinations
from typing import Dict, FrozenSet, List, Sequence


def classes(min_m: int, max_m: int, sizes: Sequence[int]) -> List[range]:
    """Piece-id range of each size class."""
    out, pid = [], 0
    for m in range(min_m, max_m + 1):
        out.append(range(pid, pid + sizes[m]))
        pid += sizes[m]
    return out


def named_sets(min_m: int, max_m: int, min_n: int, max_n: int, tgt: int,
               sizes: Sequence[int]) -> List[FrozenSet[int]]:
    """Every named set (as piece ids), in name order."""
    cls = classes(min_m, max_m, sizes)
    total = sum(len(c) for c in cls)
    size_of = [0] * total
    for d, c in enumerate(cls):
        for pid in c:
            size_of[pid] = min_m + d
    keyed = []
    for mask in range(1 << total):
        ids = [p for p in range(total) if mask >> p & 1]
        if not min_n <= len(ids) <= max_n or sum(size_of[p] for p in ids) != tgt:
            continue
        per = [tuple(p - c.start for p in ids if p in c) for c in cls]
        keyed.append((tuple(len(t) for t in per), tuple(per), frozenset(ids)))
    keyed.sort(key=lambda x: (x[0], x[1]))
    return [x[2] for x in keyed]


def name_table(min_m, max_m, min_n, max_n, tgt, sizes) -> Dict[FrozenSet[int], int]:
    return {s: i for i, s in enumerate(named_sets(min_m, max_m, min_n, max_n, tgt, sizes))}


def count_sets(min_m, max_m, min_n, max_n, tgt, sizes) -> int:
    """Number of named sets by a generating function: the coefficient of
    x^tgt y^n, n in [min_n, max_n], in prod over pieces of (1 + x^m y)."""
    # poly[c][n] = number of subsets of the pieces so far with c cells, n pieces
    poly = [[0] * (max_n + 1) for _ in range(tgt + 1)]
    poly[0][0] = 1
    for m in range(min_m, max_m + 1):
        for _ in range(sizes[m]):
            for c in range(tgt, m - 1, -1):
                for n in range(max_n, 0, -1):
                    poly[c][n] += poly[c - m][n - 1]
    return sum(poly[tgt][n] for n in range(min_n, max_n + 1))


def combination_ranks(size: int, k: int) -> Dict[tuple, int]:
    """Rank of each k-subset of range(size) in itertools.combinations order
    (lexicographic)."""
    return {c: i for i, c in enumerate(combinations(range(size), k))}
