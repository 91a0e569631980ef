"""Piece-set names (d4_naming_*, restating src/naming.hpp:15-71,
naming.inl:6-72, naming.cpp:9-81), the polyomino tables they name, and the
ledgers names are logged in (d4_ledger_*, src/search.cpp:25-226).  Host
code only: no device needed."""
import json
import os
import random
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import naming_oracle as NO  # noqa: E402


def sizes_list(max_m, by_m):
    return [by_m.get(m, 0) for m in range(max_m + 1)]


def make(d4, min_m, max_m, min_n, max_n, tgt, by_m):
    arr = [list(range(100 * m, 100 * m + by_m.get(m, 0))) for m in range(max_m + 1)]
    return d4.Naming(min_m, max_m, min_n, max_n, tgt, arr)


def all_sets(nm):
    out = []
    for v in range(nm.size()):
        u = nm.resolve_used(v)
        out.append(frozenset(i for i, x in enumerate(u) if x))
    return out


def test_reference_test_case(d4):
    """naming.cpp:83-117: sizes 2..5 with 3, 0, 7, 6 pieces, 2..8 pieces of
    20 cells: every name resolves to a set of 20 cells naming back to it."""
    by_m = {2: 3, 4: 7, 5: 6}
    nm = make(d4, 2, 5, 2, 8, 20, by_m)
    assert nm.size_pieces() == 16
    assert nm.size() == NO.count_sets(2, 5, 2, 8, 20, sizes_list(5, by_m)) > 1000
    for v in range(nm.size()):
        got = []
        assert nm.resolve(v, lambda m, i: got.append((m, i)))
        assert sum(m for m, _ in got) == 20
        assert 2 <= len(got) <= 8
        assert nm.name(lambda m, i: (m, i) in got) == v
    # and the whole order equals the brute-force listing
    assert all_sets(nm) == NO.named_sets(2, 5, 2, 8, 20, sizes_list(5, by_m))


def test_one_size_class_is_lexicographic(d4):
    """With one size class the names are the k-subsets in
    itertools.combinations order."""
    for s, k in [(7, 3), (9, 4), (5, 5), (6, 1)]:
        nm = make(d4, 4, 4, 0, 64, 4 * k, {4: s})
        ranks = NO.combination_ranks(s, k)
        assert nm.size() == len(ranks)
        for c, r in ranks.items():
            assert nm.name_used([1 if i in c else 0 for i in range(s)]) == r


def test_random_namings_match_oracle(d4):
    rng = random.Random(11)
    for _ in range(40):
        min_m = rng.randint(1, 3)
        max_m = rng.randint(min_m, 6)
        by_m = {m: rng.randint(0, 4) for m in range(min_m, max_m + 1)}
        if sum(by_m.values()) > 14:
            continue
        sizes = sizes_list(max_m, by_m)
        min_n = rng.randint(0, 3)
        max_n = rng.randint(min_n, 8)
        tgt = rng.randint(0, 24)
        nm = make(d4, min_m, max_m, min_n, max_n, tgt, by_m)
        want = NO.named_sets(min_m, max_m, min_n, max_n, tgt, sizes)
        assert nm.size() == len(want) == NO.count_sets(min_m, max_m, min_n, max_n, tgt, sizes)
        assert all_sets(nm) == want
        P = nm.size_pieces()
        for s, v in NO.name_table(min_m, max_m, min_n, max_n, tgt, sizes).items():
            assert nm.name_used([1 if p in s else 0 for p in range(P)]) == v


def test_piece_ids_and_rejections(d4):
    nm = make(d4, 2, 5, 2, 8, 20, {2: 3, 4: 7, 5: 6})
    ids = [nm.name_piece(m, i) for m in (2, 3, 4, 5) for i in range(len(nm.arr[m]))]
    assert ids == list(range(16))
    for pid in range(16):
        m, i = nm.resolve_piece(pid)
        assert nm.name_piece(m, i) == pid
    assert nm.resolve_piece(16) == (0, 0)              # out of range, as the reference
    assert nm.name_piece(3, 0) == d4.NO_NAME           # no pieces of size 3
    assert nm.resolve(nm.size(), lambda m, i: None) is False
    assert nm.name(lambda m, i: False) is None         # 0 pieces: not a named set
    assert nm.name(lambda m, i: m == 2) is None        # 3 dominoes = 6 cells != 20
    assert nm.pieces() == [200, 201, 202] + list(range(400, 407)) + list(range(500, 506))


def test_empty_and_edge_namings(d4):
    # tgt 0 with min_n 0: the empty set is the only name
    nm = make(d4, 1, 2, 0, 3, 0, {1: 2, 2: 2})
    assert nm.size() == 1 and nm.resolve_used(0) == [0, 0, 0, 0]
    assert make(d4, 1, 2, 1, 3, 0, {1: 2, 2: 2}).size() == 0
    with pytest.raises(d4.D4Error):
        make(d4, 0, 2, 0, 3, 1, {1: 1})                  # sizes start at 1
    # counts past 2^64 are refused instead of wrapping
    with pytest.raises(d4.D4Error):
        d4.Naming(1, 1, 0, 200, 100, [[], list(range(200))])


def test_polyomino_tables(d4):
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "known_shapes.json")))
    for n in range(1, 9):
        assert d4.polyominoes(n, 0xFF) == golden["free"][str(n)]
        assert d4.polyominoes(n, 0x69) == golden["one_sided"][str(n)]
        assert len(golden["one_sided"][str(n)]) == golden["one_sided_counts"][n]


# --------------------------------------------------------------------------
# ledgers
# --------------------------------------------------------------------------

def read_u64(path, off, n=1):
    import numpy as np
    return np.fromfile(path, dtype="<u8", count=n, offset=off).tolist()


def test_file_byte_layout(d4, tmp_path):
    p = str(tmp_path / "fB.bin")
    C, V = 10, 5
    L = d4.Ledger(d4.LEDGER_BYTE, C, V, p)
    assert os.path.getsize(p) == C * 32 + V * 8 + V * C
    assert L.should_run(3, 0xABC)
    assert L.log(3, 2) and not L.log(3, 2) and L.log(4, 2) and L.log(3, 0)
    assert not L.log(3, d4.NO_NAME)
    assert L.log_many([1, 1, 3], [4, 4, 2]) == [True, False, False]
    L.after_run(3, 0xABC, 2, 0)
    assert L.progress() == (4, 1)
    assert L.names() == [1, 0, 2, 0, 1]
    L.close()
    data = open(p, "rb").read()
    assert read_u64(p, 3 * 32, 4) == [0xABC, 2, 1, 0]   # ms 0 is stored as 1 (solved)
    assert read_u64(p, C * 32, V) == [1, 0, 2, 0, 1]
    led = data[C * 32 + V * 8:]
    assert led[2 * C + 3] == 0xFF and led[2 * C + 4] == 0xFF and led[0 * C + 3] == 0xFF
    assert led[4 * C + 1] == 0xFF and sum(1 for b in led if b) == 4
    # reopened: the solved config is skipped, the others run
    L = d4.Ledger(d4.LEDGER_BYTE, C, V, p)
    assert not L.should_run(3, 0xABC) and L.should_run(4, 0xDEF)
    assert not L.log(3, 2)                               # already in the file
    L.close()


def test_file_bit_layout(d4, tmp_path):
    p = str(tmp_path / "fb.bin")
    C, V = 10, 3
    L = d4.Ledger(d4.LEDGER_BIT, C, V, p)
    row = (C + 7) // 8
    assert os.path.getsize(p) == C * 32 + V * 8 + V * row
    assert L.log(9, 1) and L.log(0, 1) and not L.log(9, 1) and L.log(7, 2)
    L.close()
    led = open(p, "rb").read()[C * 32 + V * 8:]
    assert led[1 * row + 1] == 0b10 and led[1 * row + 0] == 0b1 and led[2 * row] == 0x80
    assert read_u64(p, C * 32, V) == [0, 2, 1]


def test_file_size_checks(d4, tmp_path):
    p = str(tmp_path / "x.bin")
    with open(p, "wb") as f:
        f.write(b"\0" * 7)
    with pytest.raises(d4.D4Error, match="short"):
        d4.Ledger(d4.LEDGER_BYTE, 10, 5, p)
    with open(p, "wb") as f:
        f.write(b"\0" * 10_000)
    with pytest.raises(d4.D4Error, match="long"):
        d4.Ledger(d4.LEDGER_BIT, 10, 5, p)
    with pytest.raises(d4.D4Error):
        d4.Ledger(d4.LEDGER_BYTE, 10, 5, None)


def test_atomic_count_and_verify(d4):
    L = d4.Ledger(d4.LEDGER_COUNT, 4, 3)
    assert L.log_many([0, 0, 1, 2, 2, 3], [1, 1, 1, 0, 2, 1]) == [True, False, True, True, True, True]
    assert L.names() == [1, 3, 1]
    assert L.should_run(2, 0)
    V = d4.Ledger(d4.LEDGER_VERIFY, 4, 3)
    assert V.log(0, 1) and V.log(0, 1) and not V.log(0, d4.NO_NAME)
    assert V.progress() == (2, 0)
    with pytest.raises(d4.D4Error):
        V.names()
    with pytest.raises(d4.D4Error):
        L.log(4, 0)                                      # config out of range
