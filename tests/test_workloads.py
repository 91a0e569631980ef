"""Instance producers of the reference's own workloads (src/main.cpp:14-42,
src/board.hpp:15-57), checked on the CPU with the oracle and known answers."""
import workloads


def test_calendar_puzzle_known_answers(oracle):
    """The calendar puzzle (A-Puzzle-A-Day): every real date has a solution,
    the fewest for any cell pair is 7 and the most 216 (January 25), as
    published for this puzzle; the board has 43 cells and the pieces 41."""
    shapes, counts, boards, width = workloads.calendar_boards()
    assert width == 7 and len(boards) == 12 * 31
    assert sum(bin(s).count("1") for s in shapes) == 41
    full = workloads.transform8(workloads.CALENDAR_BOARD, 3)
    assert full == oracle.transform(workloads.CALENDAR_BOARD, 3)
    assert bin(full).count("1") == 43
    got = {(m, d): oracle.count(shapes, counts, rows) for m, d, rows in boards}
    days = [31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    assert all(got[(m, d)] >= 1 for m in range(1, 13) for d in range(1, days[m - 1] + 1))
    assert min(got.values()) == 7 and max(got.values()) == 216
    assert got[(1, 25)] == 216


def test_board_configs_region_format():
    """One configuration per choice of one removed cell in every region;
    the fixed '#' cells belong to every region's shape as in
    src/board.hpp:24-30."""
    count, configs = workloads.board_configs("aa.\nbbb\n..c")
    assert count == 2 * 3 * 1 and len(configs) == 6
    # base: a a . / b b b / . . c ; every config keeps one a and two b
    for rows in configs:
        cells = sum(bin(r).count("1") for r in rows)
        assert cells == 3
        assert not (len(rows) > 2 and rows[2])  # the single c is always removed
    assert configs[0] == [0b010, 0b110]
    count, configs = workloads.board_configs("ab\n.#")
    assert count == 4  # each region also holds the '#' cell


def test_solution_map_and_steps(oracle):
    """Solution / Step expansion (src/Piece.hpp:31-48, src/Piece.cpp:25-54):
    the map covers exactly the board's open cells, each step's placed shape
    is a transform of its piece of the right size, and the matrices are
    make_step's (oracle: or_step_matrix)."""
    import ctypes
    import d4pieces
    shapes, counts, boards, width = workloads.calendar_boards()
    m, d, rows = boards[24]  # January 25
    sols = oracle.solve(shapes, counts, rows)
    assert len(sols) == 216
    for raw in sols[::27]:
        s = d4pieces.expand_solution(shapes, raw, len(rows))
        cells = {(r, c) for r, v in enumerate(rows) for c in range(8) if v >> c & 1}
        covered = {(r, c) for r in range(len(rows)) for c in range(8) if s.map[r][c] >= 0}
        assert covered == cells
        for i, st in enumerate(s.steps):
            assert bin(st.shape).count("1") == bin(shapes[st.piece_id]).count("1")
            assert d4pieces.normalize8(st.shape) == d4pieces.transform8(
                shapes[st.piece_id], st.trs_id)
            mat = (ctypes.c_int * 4)()
            oracle.L.or_step_matrix(st.trs_id, mat)
            assert (st.a, st.b, st.c, st.d) == tuple(mat)
            assert all(s.map[r][c] == i for r in range(len(rows)) for c in range(8)
                       if st.shape >> (8 * r + c) & 1)
