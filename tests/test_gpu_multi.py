"""Multi-library batches (d4_mbatch_*): every instance with its own piece set
and board width in one frontier, against the oracle (restated
src/Piece.cpp solve_count), the single-library path, and board symmetries."""
import json
import os
import random

import pytest

import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rect(w, h, blocked=()):
    rows = [(1 << w) - 1] * h
    for r, c in blocked:
        rows[r] &= 0xFF ^ (1 << c)
    return rows


def mirror(rows, w):
    """Left-right reflection within width w."""
    return [int(format(r, f"0{w}b")[::-1], 2) for r in rows]


@pytest.fixture(scope="module")
def uneven(gpu):
    return workloads.uneven_instances(4096, seed=3)


@pytest.fixture(scope="module")
def uneven_fixture():
    return json.load(open(os.path.join(GOLDEN, "uneven_4096_seed3.json")))


@pytest.fixture(params=["level", "dfs"])
def mk(request, gpu):
    """Both row-step kernels of a multi-library batch: level-synchronous
    (the default) and depth-first per lane."""
    return {"level": gpu.KERNEL_LEVEL, "dfs": gpu.KERNEL_DFS}[request.param]


def run_multi(gpu, insts, prune=True, cache=None, kernel=None):
    cache = {} if cache is None else cache
    libs = []
    for shapes, counts, rows, w in insts:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = gpu.PieceLibrary(shapes, counts, 0xFF, w)
        libs.append(cache[key])
    b = gpu.MultiBatch(libs, [x[2] for x in insts], prune=prune, kernel=kernel)
    b.run()
    return b.counts(), b.stats()


def test_random_small_mixed(gpu, oracle, mk):
    """Distinct piece sets, multiplicities and widths in one batch."""
    rng = random.Random(31)
    pool = [1, 3, 7, 259] + gpu.free_polyominoes(4) + gpu.free_polyominoes(5)
    insts = []
    for _ in range(120):
        shapes = sorted(rng.sample(pool, rng.randint(2, 6)))
        counts = [rng.randint(1, 3) for _ in shapes]
        w, h = rng.randint(2, 8), rng.randint(1, 6)
        blocked = [(rng.randrange(h), rng.randrange(w)) for _ in range(rng.randint(0, 3))]
        insts.append((shapes, counts, rect(w, h, blocked), w))
    got, st = run_multi(gpu, insts, kernel=mk)
    want = [oracle.count(s, c, r) for s, c, r, _ in insts]
    assert got == want
    assert sum(want) > 0 and st["retries"] >= 0
    # pruning changes the work, never the counts
    got2, _ = run_multi(gpu, insts, prune=False, kernel=mk)
    assert got2 == want


def test_more_than_64_candidates(gpu, oracle, mk):
    """Libraries whose (piece, transform) pairs need a 128-bit mask: 14 free
    hexominoes (subset tilings of small boards)."""
    p6 = gpu.free_polyominoes(6)
    rng = random.Random(5)
    boards = [(6, 2), (6, 3), (4, 3), (8, 3), (6, 4), (4, 6), (6, 5), (3, 8), (6, 6)]
    insts = []
    for w, h in boards:
        shapes = sorted(rng.sample(p6, 14))
        insts.append((shapes, [1] * 14, rect(w, h), w))
    libs = [gpu.PieceLibrary(s, c, 0xFF, w) for s, c, _, w in insts]
    assert max(L.n_cand for L in libs) > 64
    got, _ = run_multi(gpu, insts, kernel=mk)
    want = [oracle.count(s, c, r) for s, c, r, _ in insts]
    assert got == want
    assert sum(1 for x in want if x) >= 3
    # the single-library searcher refuses such a library with a clear error
    big = next(L for L in libs if L.n_cand > 64)
    with pytest.raises(gpu.D4Error):
        gpu.CudaSearcher(big, rows=rect(big.width, 2))


def test_single_library_batch_agrees(gpu, mk):
    """The 256 seeded 8x8 boards through the multi-library kernel equal the
    oracle's counts (fixture written by tests/golden/make_seeded_counts.py)."""
    fx = json.load(open(os.path.join(GOLDEN, "seeded_8x8_seed2.json")))
    p5 = gpu.free_polyominoes(5)
    insts = [(p5, [1] * 12, rows, 8) for rows in fx["boards"]]
    got, st = run_multi(gpu, insts, kernel=mk)
    assert got == fx["counts"]
    assert st["rows_done"] == 9


def test_memory_cap_splits_frontier(gpu, mk):
    """A memory cap far below the frontier's size parks frontier halves and
    finishes them one after the other: counts are unchanged."""
    fx = json.load(open(os.path.join(GOLDEN, "seeded_8x8_seed2.json")))
    p5 = gpu.free_polyominoes(5)
    lib = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 8)
    n = 16
    b = gpu.MultiBatch([lib] * n, fx["boards"][:n], kernel=mk)
    b.set_memory(24 << 20)
    b.run()
    assert b.counts() == fx["counts"][:n]
    assert b.stats()["splits"] > 0


def test_budget_drops_instances(gpu, mk):
    """Instances whose frontier share exceeds the state budget leave the
    search and are reported unfinished; the others keep exact counts."""
    fx = json.load(open(os.path.join(GOLDEN, "seeded_8x8_seed2.json")))
    p5 = gpu.free_polyominoes(5)
    lib = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 8)
    n = 32
    b = gpu.MultiBatch([lib] * n, fx["boards"][:n], kernel=mk)
    b.run()
    peak, fin = b.instances()
    assert all(fin) and b.counts() == fx["counts"][:n]
    budget = sorted(peak)[n // 2]
    b2 = gpu.MultiBatch([lib] * n, fx["boards"][:n], kernel=mk)
    b2.set_budget(budget)
    b2.run()
    peak2, fin2 = b2.instances()
    got = b2.counts()
    assert 0 < sum(fin2) < n
    for i in range(n):
        if fin2[i]:
            assert got[i] == fx["counts"][i] and peak2[i] <= budget
        else:
            assert got[i] == 0 and peak2[i] > budget


def test_tune_keeps_counts(gpu, oracle):
    """The multi-batch autotuner picks a launch shape from its sample; the
    counts do not depend on it."""
    rng = random.Random(8)
    pool = [3, 7, 259] + gpu.free_polyominoes(4)
    insts = []
    for _ in range(40):
        shapes = sorted(rng.sample(pool, rng.randint(2, 5)))
        counts = [rng.randint(1, 2) for _ in shapes]
        w, h = rng.randint(3, 7), rng.randint(2, 6)
        insts.append((shapes, counts, rect(w, h), w))
    libs = [gpu.PieceLibrary(s, c, 0xFF, w) for s, c, _, w in insts]
    b = gpu.MultiBatch(libs, [x[2] for x in insts])
    t = b.tune(16)
    assert t["threads"] in (256, 512, 1024) and t["refill"] in (8, 16, 32)
    b.run()
    assert b.counts() == [oracle.count(s, c, r) for s, c, r, _ in insts]


def test_edge_cases(gpu, oracle, mk):
    p4 = gpu.free_polyominoes(4)
    insts = [
        ([3], [1], [], 2),                 # empty board: one empty tiling
        ([3], [1], [0, 0], 2),             # all cells blocked
        ([3], [0], [3], 2),                # piece with multiplicity 0
        ([7], [4], rect(3, 4), 3),         # one piece kind, several copies
        (p4, [1] * 5, rect(4, 5), 4),      # 5 tetrominoes on 4x5
        ([1], [64], rect(8, 8), 8),        # monominoes, 7-bit counter
    ]
    got, _ = run_multi(gpu, insts, kernel=mk)
    want = [oracle.count(s, c, r) for s, c, r, _ in insts]
    assert want[0] == 1 and want[1] == 1 and want[5] == 1
    assert got == want


def test_uneven_instances(gpu, uneven, uneven_fixture, mk):
    """The uneven puzzles (first 1024 of the 4096) in one batch with a state
    budget of 2^20 per instance: every instance that finishes and that the
    oracle finished within its node budget (fixture written by
    tests/golden/make_uneven_counts.py) has the oracle's count; the count is
    invariant under the board's left-right reflection (same frontier sizes,
    and top-bottom reflection; the piece sets are closed under D4.  (Which
    instances finish can differ between the runs when the frontier is split
    to fit in memory, since the budget applies per frontier part: counts are
    compared where both runs finish.)"""
    from golden.make_uneven_counts import digest
    assert digest(uneven) == uneven_fixture["sha256"]
    n, budget = 1024, 1 << 20
    insts = uneven[:n]
    cache = {}

    def run(batch):
        libs = []
        for shapes, counts, _, w in batch:
            key = (tuple(shapes), tuple(counts), w)
            if key not in cache:
                cache[key] = gpu.PieceLibrary(shapes, counts, 0xFF, w)
            libs.append(cache[key])
        b = gpu.MultiBatch(libs, [x[2] for x in batch], kernel=mk)
        b.set_budget(budget)
        b.run()
        return b.counts(), b.instances()[1]

    got, fin = run(insts)
    pinned = {int(k): v for k, v in uneven_fixture["counts"].items() if int(k) < n}
    checked = [i for i in pinned if fin[i]]
    assert len(checked) >= 300
    bad = [(i, got[i], pinned[i]) for i in checked if got[i] != pinned[i]]
    assert not bad, bad[:10]
    assert sum(fin) >= n // 2
    assert all(got[i] == 0 for i in range(n) if not fin[i])
    for reflect in (lambda r, w: mirror(r, w), lambda r, w: r[::-1]):
        got_r, fin_r = run([(s, c, reflect(r, w), w) for s, c, r, w in insts])
        both = [i for i in range(n) if fin[i] and fin_r[i]]
        assert len(both) >= n // 2
        bad = [(i, got[i], got_r[i]) for i in both if got_r[i] != got[i]]
        assert not bad, bad[:10]


def test_donation_buffer_full(gpu, uneven, uneven_fixture):
    """Regression for the donation-buffer-full branch of k_extend_multi (the
    path whose counter once wrapped past 2^32): a 64-item buffer and an
    8-step budget make lanes find it full many times; counts still equal the
    oracle's where instances finish, and the refusals are counted."""
    n, budget = 256, 1 << 20
    insts = uneven[:n]
    cache, libs = {}, []
    for shapes, counts, _, w in insts:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = gpu.PieceLibrary(shapes, counts, 0xFF, w)
        libs.append(cache[key])
    b = gpu.MultiBatch(libs, [x[2] for x in insts], kernel=gpu.KERNEL_DFS)
    b.set_budget(budget)
    b.set_budget_steps(8)
    b.set_donation_capacity(64)
    b.run()
    got, fin = b.counts(), b.instances()[1]
    st = b.stats()
    assert st["donation_full"] > 0 and st["donated"] > 0
    pinned = {int(k): v for k, v in uneven_fixture["counts"].items() if int(k) < n}
    checked = [i for i in pinned if fin[i]]
    assert len(checked) > 100
    assert all(got[i] == pinned[i] for i in checked)


def test_device_tables_equal_host_tables(gpu, uneven, oracle):
    """The per-library tables a batch builds on the device equal, byte for
    byte, the host build of the same libraries: all distinct libraries of the
    4096 uneven puzzles (sizes 3-6, up to 14 kinds, widths 4-8), libraries
    with more than 64 placements, one-sided pieces, zero counts and
    multiplicities up to 3.  Batches built either way count alike, and like
    the oracle on small boards."""
    cache, libs, boards = {}, [], []
    for shapes, counts, rows, w in uneven:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = gpu.PieceLibrary(shapes, counts, 0xFF, w)
            libs.append(cache[key])
            boards.append(rows)
    rng = random.Random(5)
    pool = gpu.free_polyominoes(3) + gpu.free_polyominoes(4) + gpu.free_polyominoes(5)
    extra = []
    for sym in (0xFF, 0x69, 0x01):
        for _ in range(20):
            shapes = sorted(rng.sample(pool, rng.randint(1, 9)))
            counts = [rng.randint(0, 3) for _ in shapes]
            w = rng.randint(3, 8)
            lib = gpu.PieceLibrary(shapes, counts, sym, w)
            libs.append(lib)
            boards.append(rect(w, 2))
            extra.append((shapes, counts, sym, w, lib))
    b = gpu.MultiBatch(libs, boards)
    for i, lib in enumerate(libs):
        assert b.tables(i) == lib.host_tables(), i
    bh = gpu.MultiBatch(libs, boards, host_tables=True)
    for i in range(0, len(libs), 97):
        assert bh.tables(i) == b.tables(i)
    small = [(s, c, sym, w, lib) for s, c, sym, w, lib in extra if sym == 0xFF][:12]
    sb = [rect(w, 3) for _, _, _, w, _ in small]
    bd = gpu.MultiBatch([x[4] for x in small], sb)
    bd.run()
    bh2 = gpu.MultiBatch([x[4] for x in small], sb, host_tables=True)
    bh2.run()
    assert bd.counts() == bh2.counts()
    for (s, c, _, w, _), rows, n in zip(small, sb, bd.counts()):
        assert n == oracle.count(s, c, rows)


def test_kernels_agree_on_work(gpu, uneven, uneven_fixture):
    """The level and depth-first kernels walk the same search: on 512 uneven
    puzzles with a state budget, the same instances finish with the same
    counts (the oracle's where it has them), and the rows extended, children
    kept, children pruned and unique states match exactly, although the
    level kernels rerun the chunks whose buffers overflow (the first row's
    2 Mi children exceed the initial room).  The expand kernel is refused."""
    n, budget = 512, 1 << 20
    insts = uneven[:n]
    cache, libs = {}, []
    for shapes, counts, _, w in insts:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = gpu.PieceLibrary(shapes, counts, 0xFF, w)
        libs.append(cache[key])
    res = {}
    for k in (gpu.KERNEL_LEVEL, gpu.KERNEL_DFS):
        b = gpu.MultiBatch(libs, [x[2] for x in insts], kernel=k)
        b.set_budget(budget)
        b.run()
        res[k] = (b.counts(), b.instances()[1], b.stats())
        b = None
    (cl, fl, sl), (cd, fd, sd) = res[gpu.KERNEL_LEVEL], res[gpu.KERNEL_DFS]
    assert fl == fd and cl == cd
    assert sl["splits"] == 0 and sd["splits"] == 0
    for f in ("records_extended", "children", "pruned", "unique", "rows_done"):
        assert sl[f] == sd[f], f
    assert sl["donated"] == 0 and sl["chunk_reruns"] > 0 and sd["chunk_reruns"] == 0
    pinned = {int(k): v for k, v in uneven_fixture["counts"].items() if int(k) < n}
    checked = [i for i in pinned if fl[i]]
    assert len(checked) >= 150 and all(cl[i] == pinned[i] for i in checked)
    b = gpu.MultiBatch(libs[:4], [x[2] for x in insts[:4]])
    with pytest.raises(gpu.D4Error):
        b.set_kernel(gpu.KERNEL_EXPAND)
