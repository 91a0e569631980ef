"""The consumer of terminal states on the device: every configuration of a
board file is tiled with a Naming's pieces (each at most once) in one
multi-instance frontier; the distinct piece sets that reach a terminal state
are named and logged (src/search.cpp:25-320, finishing what
src/searcher.cpp:38-65 leaves commented out).  Checked against the oracle's
enumeration of every tiling (src/Piece.cpp:66-102 restated)."""
import os
from collections import defaultdict

import pytest

import searchers
import workloads

pytestmark = pytest.mark.gpu

# 20 cells in two regions: one cell of each is removed per configuration
BOARD = """
aaaaa
aaaab
bbbbb
bbbbb
"""


def oracle_sets(oracle, pieces, rows, sym, min_n, max_n):
    """{frozenset(piece ids): tilings} over every tiling of the board."""
    out = defaultdict(int)
    for sol in oracle.solve(pieces, [1] * len(pieces), rows, sym=sym):
        ids = frozenset(s[0] for s in sol)
        assert len(ids) == len(sol)
        if min_n <= len(ids) <= max_n:
            out[ids] += 1
    return out


def resolve_set(nm, v):
    return frozenset(i for i, u in enumerate(nm.resolve_used(v)) if u)


@pytest.mark.parametrize("sym_C,min_m,max_m,min_n,max_n", [
    (False, 2, 5, 4, 6),   # 20 free pieces, 90 (piece, transform) placements
    (True, 2, 4, 3, 7),    # one-sided pieces (the reference's C=1 mode)
])
def test_ledgers_match_oracle(gpu, oracle, tmp_path, sym_C, min_m, max_m, min_n, max_n):
    count, configs = workloads.board_configs(BOARD)
    assert count == len(configs) == 99
    print(f"ledger case sym_C={sym_C} sizes {min_m}..{max_m} n {min_n}..{max_n}", flush=True)
    path = str(tmp_path / "fB.bin")
    fB, run = searchers.main(BOARD, min_m, max_m, min_n, max_n, "fB", path=path, sym_C=sym_C)
    nm, pieces, sym = fB.naming, fB.naming.pieces(), fB.library.sym
    assert run["configs"] == 99
    want = [oracle_sets(oracle, pieces, rows, sym, min_n, max_n) for rows in configs]
    # every (set, config) pair the file holds, against the oracle's sets
    C, V = len(configs), nm.size()
    data = open(path, "rb").read()
    led = data[C * 32 + V * 8:]
    got = [set() for _ in range(C)]
    for v in range(V):
        row = led[v * C:(v + 1) * C]
        for c in range(C):
            if row[c]:
                assert row[c] == 0xFF
                got[c].add(resolve_set(nm, v))
    for c in range(C):
        assert got[c] == set(want[c]), c
    assert sum(len(w) for w in want) > 500
    assert fB.zero == [c for c in range(C) if not want[c]]
    # AtomicCount's per-name totals equal the file's name header, FileBit's
    # bits the file's bytes, Verify's work the number of pairs
    ac, _ = searchers.main(BOARD, min_m, max_m, min_n, max_n, "ac", sym_C=sym_C)
    import numpy as np
    nmes = np.frombuffer(data[C * 32:C * 32 + V * 8], dtype="<u8").tolist()
    assert ac.ledger.names() == nmes == fB.ledger.names()
    fb, _ = searchers.main(BOARD, min_m, max_m, min_n, max_n, "fb",
                           path=str(tmp_path / "fb.bin"), sym_C=sym_C)
    bits = open(str(tmp_path / "fb.bin"), "rb").read()[C * 32 + V * 8:]
    rb = (C + 7) // 8
    for v in range(V):
        for c in range(C):
            assert bool(bits[v * rb + c // 8] >> (c % 8) & 1) == bool(led[v * C + c])
    vf, _ = searchers.main(BOARD, min_m, max_m, min_n, max_n, "v", sym_C=sym_C)
    assert vf.work_done() == sum(len(w) for w in want) == fB.work_done()
    # the descriptors: shape and count per config, all solved
    desc = np.frombuffer(data[:C * 32], dtype="<u8").reshape(C, 4)
    assert desc[:, 0].tolist() == [searchers.shape_of(r) for r in configs]
    assert desc[:, 1].tolist() == [len(w) for w in want]
    assert (desc[:, 2] >= 1).all()
    # a second run over the same file finds every config solved
    fB.ledger.close()
    again, run2 = searchers.main(BOARD, min_m, max_m, min_n, max_n, "fB", path=path, sym_C=sym_C)
    assert run2["configs"] == 0 and again.ledger.names() == nmes


def test_terminal_states_and_tilings(gpu, oracle):
    """d4_mbatch_terminals: per (config, piece set) the number of tilings
    ending in it equals the oracle's; totals equal the batch's counts."""
    count, configs = workloads.board_configs(BOARD)
    nm = gpu.Naming(2, 5, 0, 64, 18, searchers.known_shapes())
    pieces = nm.pieces()
    lib = gpu.PieceLibrary(pieces, [1] * len(pieces), 0xFF, 8)
    sub = configs[::7]
    b = gpu.MultiBatch([lib] * len(sub), sub)
    b.set_terminals(True)
    b.run()
    inst, ctr, til = b.terminals()
    assert list(zip(inst, ctr)) == sorted(set(zip(inst, ctr)))
    per = defaultdict(dict)
    for i, c, t in zip(inst, ctr, til):
        per[i][resolve_set(nm, nm.name_counters(lib, [c])[0])] = t
    for i, rows in enumerate(sub):
        assert per[i] == dict(oracle_sets(oracle, pieces, rows, 0xFF, 0, 64)), i
    assert b.counts() == [sum(per[i].values()) for i in range(len(sub))]


def test_terminal_sink_grows_and_splits(gpu, oracle):
    """A memory cap that splits the frontier, and more terminal states than
    the sink first holds: the states are unchanged."""
    count, configs = workloads.board_configs(BOARD)
    nm = gpu.Naming(2, 5, 0, 64, 18, searchers.known_shapes())
    pieces = nm.pieces()
    lib = gpu.PieceLibrary(pieces, [1] * len(pieces), 0xFF, 8)
    b = gpu.MultiBatch([lib] * len(configs), configs)
    b.set_terminals(True, 1 << 20)
    b.run()
    ref = b.terminals()
    assert b.stats()["sink_growths"] == 0 and len(ref[0]) > 30000
    b2 = gpu.MultiBatch([lib] * len(configs), configs)
    b2.set_terminals(True, 64)
    b2.set_memory(2 << 20)
    b2.run()
    st = b2.stats()
    assert st["splits"] > 0 and st["sink_growths"] > 0
    assert b2.terminals() == ref and b2.counts() == b.counts()
