/*
 * oracle.h -- CPU restatement of the reference's polyomino backtracker.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library under
 * d4-pieces_amd/ and its C ABI in include/d4.h) links, loads or calls this
 * code.  It is used by tests/ as the parity checker, by
 * __graft_entry__.smoke() as the checker of one small GPU run, and by
 * bench.py's cpu_baseline leg as the timed CPU baseline ("port").
 *
 * What it restates (reference files, read as text):
 *   - Shape<L>::transform / transforms / normalize   src/Shape.cpp:6-40,
 *     src/Shape.hpp:184-195   (8 D4 transforms, reference order)
 *   - Piece<L>::Piece (unique transforms per sym mask) src/Piece.cpp:12-22
 *   - Piece<L>::cover (anchor = lowest open cell)     src/Piece.inl:5-24
 *   - min_tiles / solve / solve_count                 src/Piece.cpp:57-139
 *   - make_step (2x2 matrix per transform)            src/Piece.cpp:41-54
 *   - canonical_form (min over transforms)            src/Shape.cpp:57-66
 * generalised from the reference's square Shape<8>/Shape<11> bitboards to a
 * board of W <= 8 columns (row stride 8 bits) and H <= 64 rows, because the
 * 6x10, 8x15 and tall batched boards do not fit the reference's types.  The
 * DFS order, pruning and multiplicity semantics are unchanged.
 *
 * Pinning (which checks fix each function; see tests/test_oracle.py):
 *   or_transform      -> group-algebra closure + symmetry classes of the
 *                        pentominoes/tetrominoes (reference Group.cpp asserts),
 *                        the 63 fixed pentominoes (OEIS A001168)
 *   or_enum_free      -> OEIS A000105 counts and the reference's published
 *                        known1.cpp shape values (golden fixture)
 *   or_solve_count    -> 6x10/5x12/4x15/3x20 pentomino KATs, 8x8-minus-centre
 *                        520, domino Fibonacci and Kasteleyn 8x8 12,988,816,
 *                        straight-tromino recurrence, brute force on tiny boards
 *   or_solve          -> count equality with or_solve_count, exact-cover and
 *                        multiplicity invariants, D2 symmetry closure of the
 *                        6x10 solution set, first-solution = lexicographic
 *                        minimum of the full set
 *   or_parse_R /      -> round trip parse(assemble(x)) == x for random records
 *   or_assemble_R        at every height (the reference's bmark/record.cu
 *                        check), and hand-computed byte layouts of record.h
 */
#ifndef D4_ORACLE_H
#define D4_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* D4 transform t (0..7, reference order: id, flipX, flipY, rot180,
 * transpose, rot90 one way, rot90 the other, anti-transpose) of a shape in
 * the 8-wide encoding (bit = row*8 + col, shape within 8x8), normalised to
 * the top-left corner.  Shape.cpp:6-40. */
uint64_t or_transform(uint64_t shape, unsigned t);

/* Canonical form = minimum over the transforms enabled in `forms` of the
 * normalised transform.  Shape.cpp:57-66. */
uint64_t or_canonical(uint64_t shape, unsigned forms);

/* Enumerate every free polyomino with n cells (1 <= n <= 8) as its canonical
 * value; writes them sorted ascending into out[0..cap) and returns the total
 * number.  (The reference ships these as tables, known1.cpp; enum.cpp makes
 * them by brute force.) */
size_t or_enum_free(unsigned n, uint64_t *out, size_t cap);

/* Unique transforms of a piece under sym mask: writes up to 8 normalised
 * shapes and their transform indices (reference order, duplicates dropped
 * as in Piece.cpp:12-22); returns the number written. */
unsigned or_piece_orients(uint64_t shape, unsigned sym, uint64_t *normals,
                          unsigned *trs);

/* solve_count (Piece.cpp:105-139) on a board of H rows (row stride 8 bits,
 * rows[r] bit c = cell (r, c) is open).  shapes[i] / counts[i] = piece i's
 * shape and multiplicity, tried in library order i = 0..n-1; sym = enabled
 * transform mask (0xff = D4).  Returns the number of tilings. */
uint64_t or_solve_count(const uint64_t *shapes, const uint32_t *counts,
                        uint32_t n, uint32_t sym, const uint8_t *rows,
                        uint32_t H);

/* or_solve_count with a budget of DFS nodes (calls of the recursion):
 * *complete = 0 when the budget ran out (the count is then partial). */
uint64_t or_solve_count_budget(const uint64_t *shapes, const uint32_t *counts,
                               uint32_t n, uint32_t sym, const uint8_t *rows,
                               uint32_t H, uint64_t max_nodes, int *complete);

/* solve (Piece.cpp:66-102).  Each solution is written to `out` as
 *   u8 nsteps, then nsteps x {u8 piece_id, u8 trs_id, u8 y, u8 x}
 * in the DFS order (= anchor order), where (y, x) is the translation of the
 * piece's normalised transform (Step::y, Step::x).  single != 0 stops at the
 * first solution.  Returns the number of solutions found; *out_len receives
 * the bytes needed (if it exceeds cap, output is truncated but the count is
 * still complete). */
uint64_t or_solve(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                  uint32_t sym, const uint8_t *rows, uint32_t H, int single,
                  uint8_t *out, size_t cap, size_t *out_len);

/* or_solve with each piece's placements switched on by its own mask
 * (Piece::Placement::enabled, Piece.hpp:16-20): bit p of enabled[i] = the
 * p-th transform sym admits (its trs_id), duplicate shapes included when
 * their bit is set.  enabled == NULL gives or_solve (duplicates off). */
uint64_t or_solve_ex(const uint64_t *shapes, const uint32_t *counts, const uint8_t *enabled,
                     uint32_t n, uint32_t sym, const uint8_t *rows, uint32_t H, int single,
                     uint8_t *out, size_t cap, size_t *out_len);

/* make_step's 2x2 matrix for transform trs (Piece.cpp:41-54): a, b, c, d. */
void or_step_matrix(unsigned trs, int *abcd);

/* The reference's frontier record codec (src/record.h:5-32,
 * src/record.cuh:5-71).  r[0..4] = R {xaL, xaH, ex0, ex1, ex2}.
 *   or_parse_R:    parse_R<H>(R, ea) -> empty_area (row 0 = ea), nm_cnt,
 *                  nm[16] (ids, 0xff padded)
 *   or_assemble_R: the R part of assemble_R<H>(RCfg) for an RCfg whose
 *                  empty_area has the just-filled row at bits 0..7 (so the
 *                  result's ea is empty_area >> 8 & 0xff, returned in *ea).
 * H is the height of the record produced / parsed (1..8; assemble 1..7).
 * __byte_perm is evaluated as CUDA's prmt default mode (selector bit 3
 * replicates the selected byte's sign bit). */
void or_parse_R(unsigned H, const uint32_t *r, uint8_t ea, uint64_t *empty_area,
                uint32_t *nm_cnt, uint8_t *nm);
void or_assemble_R(unsigned H, uint64_t empty_area, uint32_t nm_cnt, const uint8_t *nm,
                   uint32_t *r, uint8_t *ea);

#ifdef __cplusplus
}
#endif
#endif
