"""The frontier exported in the reference's RX record format equals an
independent row-by-row enumeration (restated in Python from the row-step
semantics, piece orientations from the oracle) encoded with the oracle's
restatement of assemble_R (src/record.cuh:43-71)."""
import pytest

pytestmark = pytest.mark.gpu


def row_bfs(oracle, shapes, rows, upto):
    """Frontiers after rows 0..upto-1: {(open window u64, used ids): paths}."""
    cands = []
    for i, s in enumerate(shapes):
        for nm, _ in oracle.orients(s):
            mc = (nm & 0xFF & -(nm & 0xFF)).bit_length() - 1
            cands.append((i, nm, mc))
    board = list(rows) + [0] * 16

    def bw(r):
        return sum(board[r + j] << (8 * j) for j in range(8))

    front = {(bw(0), frozenset()): 1}
    out = [front]
    for r in range(upto):
        nxt = {}
        for (E, used), cnt in front.items():
            def dfs(E, used):
                if not E & 0xFF:
                    key = ((E >> 8) | (board[r + 8] << 56), used)
                    nxt[key] = nxt.get(key, 0) + cnt
                    return
                c = ((E & 0xFF) & -(E & 0xFF)).bit_length() - 1
                for i, nm, mc in cands:
                    if i in used or c < mc:
                        continue
                    pl = nm << (c - mc)
                    if (pl & 0xFF) >> c << c != pl & 0xFF or pl & ~E:
                        continue
                    if any(((nm >> (8 * j)) & 0xFF) << (c - mc) > 0xFF for j in range(8)):
                        continue
                    dfs(E & ~pl, used | {i})
            dfs(E, used)
        front = nxt
        out.append(front)
    return out


def encode(oracle, E, used, H):
    ids = sorted(used)
    if H == 8:
        return (E & 0xFF, (E >> 8) & 0xFFFFFFFF, (E >> 40) & 0xFFFFFFFF, 0, 0, 0)
    r, ea = oracle.assemble_R(H, E << 8 & ((1 << 64) - 1), len(ids), ids)
    if H == 7:
        r = r[:4] + (0xFFFFFFFF,)  # ex2 is not part of a height-7 record
    return (ea,) + tuple(r)


@pytest.mark.parametrize("width,rows,upto", [
    (8, [0xFF, 0xFD, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xF7], 2),
    (6, [0x3F, 0x3F, 0x3F, 0x3B, 0x3F], 5),
])
def test_export_matches_row_bfs(gpu, oracle, width, rows, upto):
    p5 = gpu.free_polyominoes(5)
    lib = gpu.PieceLibrary(p5, [1] * 12, 0xFF, width)
    s = gpu.CudaSearcher(lib, rows=rows, prune=False)
    fronts = row_bfs(oracle, p5, rows, upto)
    for r in range(upto + 1):
        if not s.get_height():
            break
        H = len(rows) - r
        got = [(ea,) + tuple(w) + (c,) for ea, w, c in s.export()]
        want = sorted(encode(oracle, E, used, H) + (cnt,)
                      for (E, used), cnt in fronts[r].items())
        assert got == want, r
        if r < upto:
            s.search_GPU()


def test_size_buckets(gpu):
    """size(i) / next_size(i) (src/searcher_cuda.h:16-27): the records of
    bucket i are the exported RX records with ea == i (the reference's WL
    bucket `pos`), the buckets add up to size() and next_size(), and each
    bucket's bound is the row-0 completion count of i times its records."""
    import collections
    p5 = gpu.free_polyominoes(5)
    lib = gpu.PieceLibrary(p5, [1] * 12, 0xFF, 8)
    rows = [0xFF, 0xFF, 0xFF, 0xE7, 0xE7, 0xFF, 0xFF, 0xFF]
    s = gpu.CudaSearcher(lib, rows=rows)
    per_pos = {}
    steps = 0
    while s.get_height():
        hist = collections.Counter(ea for ea, _, _ in s.export())
        sizes = [s.size(i) for i in range(256)]
        nexts = [s.next_size(i) for i in range(256)]
        assert sizes == [hist.get(i, 0) for i in range(256)]
        assert sum(sizes) == s.size()
        assert sum(nexts) == s.next_size()
        for i in range(256):
            if sizes[i]:
                per = nexts[i] // sizes[i]
                assert nexts[i] == per * sizes[i]
                assert per_pos.setdefault(i, per) == per  # depends on i alone
            else:
                assert nexts[i] == 0
        assert s.size(256) == 0
        s.search_GPU()
        steps += 1
    assert steps == 8 and per_pos[0xFF] > 1
