"""The allocation-free entry point (d4_searcher_count_noalloc) against the
allocating searcher path (d4_searcher_create + run + count) on the same
boards, through the C ABI: two calls in a row, the second on a smaller board
(the scratch keeps the first call's contents), plus the oracle's counts and
the error for a search larger than the scratch."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def rect(w, h, blocked=()):
    rows = [(1 << w) - 1] * h
    for r, c in blocked:
        rows[r] &= 0xFF ^ (1 << c)
    return rows


def searcher_count(d4, lib, rows):
    s = d4.CudaSearcher(lib, rows=rows)
    s.run()
    return s.count()


def test_noalloc_matches_searcher(gpu, oracle):
    p5 = gpu.free_polyominoes(5)
    lib8 = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 8)
    centre = rect(8, 8, [(3, 3), (3, 4), (4, 3), (4, 4)])
    big = gpu.count_noalloc(lib8, centre)
    assert big == searcher_count(gpu, lib8, centre) == 520
    # second call, smaller board and another library on the same device
    lib5 = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 5)
    small = gpu.count_noalloc(lib5, rect(5, 12))
    assert small == searcher_count(gpu, lib5, rect(5, 12)) == 4040
    # again on the first board: nothing left over from the calls before
    assert gpu.count_noalloc(lib8, centre) == 520
    # seeded boards, with and without pruning, against the oracle
    import workloads
    for rows in workloads.seeded_8x8(3, seed=2):
        want = oracle.count(p5, [1] * 12, rows)
        assert gpu.count_noalloc(lib8, rows) == want == searcher_count(gpu, lib8, rows)
    assert gpu.count_noalloc(lib5, rect(5, 12), prune=False) == 4040
    assert gpu.count_noalloc(lib8, []) == 1


def test_noalloc_rejects_oversized(gpu):
    """Doubled pentominoes on 8x15 need a frontier far beyond the scratch:
    a clear D4_ERANGE, and the entry point keeps working afterwards."""
    p5 = gpu.free_polyominoes(5)
    lib = gpu.PieceLibrary(p5, [2] * 12, 0xFF, 8)
    with pytest.raises(gpu.D4Error) as ei:
        gpu.count_noalloc(lib, rect(8, 15))
    assert ei.value.code == -5 and "allocation-free" in str(ei.value)
    lib5 = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 5)
    assert gpu.count_noalloc(lib5, rect(5, 12)) == 4040
