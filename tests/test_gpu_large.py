"""The large single enumeration: the 8x15 board with each of the 12 free
pentominoes available twice (24 pieces, 120 cells; peak frontier 3.1e9
states).  The oracle cannot finish it, so the full count is pinned by two
kernels that share no search code -- the level-synchronous row step (the
default) and the depth-first one -- and the frontier states the search
passes 3 rows before the end are each counted on the GPU and by the oracle
(restated src/Piece.cpp solve_count) on the sub-board they leave."""
import os
import sys

import pytest

import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LARGE_COUNT = 15_544_030_990_612


def test_large_count_two_kernels(gpu, oracle):
    sys.path.insert(0, ROOT)
    import bench
    shapes, counts, rows, w = workloads.large_8x15()
    assert len(rows) == 15 and w == 8 and list(counts) == [2] * 12
    counted = {}
    for kernel in (gpu.KERNEL_LEVEL, gpu.KERNEL_DFS):
        lib = gpu.PieceLibrary(shapes, counts, 0xFF, w)
        lib.set_kernel(kernel)
        s = gpu.CudaSearcher(lib, rows=rows)
        if kernel == gpu.KERNEL_LEVEL:
            s.set_sample(3, 16)
        s.run()
        counted[kernel] = (s.count(), s.stats()["peak_frontier"])
        if kernel == gpu.KERNEL_LEVEL:
            sample = s.sample()
            layout = lib.counter_layout()
        s = None
        gpu.release_cached()
    assert counted[gpu.KERNEL_LEVEL][0] == LARGE_COUNT
    assert counted[gpu.KERNEL_DFS][0] == LARGE_COUNT
    assert counted[gpu.KERNEL_LEVEL][1] == counted[gpu.KERNEL_DFS][1] > 3_000_000_000
    # states 3 rows before the end: the sub-board each leaves counts alike
    assert len(sample) == 16
    for key, cnt in sample[:8]:
        sub, left = bench.sub_board(rows, len(rows) - 3, key, lib.max_height, layout, counts)
        assert len(sub) == 3 and cnt > 0
        want = oracle.count(shapes, left, sub)
        sl = gpu.PieceLibrary(shapes, left, 0xFF, w)
        ss = gpu.CudaSearcher(sl, rows=sub)
        ss.run()
        assert ss.count() == want
