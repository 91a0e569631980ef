/*
 * oracle.c -- CPU restatement of the reference backtracker.  TEST
 * INFRASTRUCTURE ONLY (see oracle.h for what may use it and how each
 * function is pinned).  Plain C, no dependencies.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* geometry: 8-wide bitboards (bit = row*8 + col)                      */
/* ------------------------------------------------------------------ */

static int test_bit(uint64_t s, unsigned r, unsigned c) {
    return (int)((s >> (r * 8u + c)) & 1u);
}

static uint64_t normalize8(uint64_t v) {
    if (!v) return 0;
    while (!(v & 0xffull)) v >>= 8;                /* Shape::normalize rows */
    while (!(v & 0x0101010101010101ull)) v >>= 1;  /* ... then columns      */
    return v;
}

/* Shape.cpp:6-26 with LEN = 8: output cell (r, c) takes input cell
 * (Swap ? c : r, Swap ? r : c), then FlipX mirrors the column and FlipY the
 * row.  Transform index order is Shape.cpp:28-40. */
uint64_t or_transform(uint64_t s, unsigned t) {
    static const unsigned char SWAP[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    static const unsigned char FX[8]   = {0, 1, 0, 1, 0, 1, 0, 1};
    static const unsigned char FY[8]   = {0, 0, 1, 1, 0, 0, 1, 1};
    uint64_t out = 0;
    if (!s) return 0;
    for (unsigned r = 0; r < 8; r++)
        for (unsigned c = 0; c < 8; c++) {
            unsigned ir = SWAP[t & 7] ? c : r;
            unsigned ic = SWAP[t & 7] ? r : c;
            if (FX[t & 7]) ic = 7 - ic;
            if (FY[t & 7]) ir = 7 - ir;
            if (test_bit(s, ir, ic)) out |= 1ull << (r * 8 + c);
        }
    return normalize8(out);
}

uint64_t or_canonical(uint64_t s, unsigned forms) {
    uint64_t v = ~0ull;
    for (unsigned t = 0; t < 8; t++, forms >>= 1)
        if (forms & 1u) {
            uint64_t x = or_transform(s, t);
            if (x < v) v = x;
        }
    return v;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* Free polyominoes of size n: grow every size n-1 free polyomino by one
 * 4-neighbour cell and keep canonical forms.  (enum.cpp brute-forces
 * bitmasks instead; the resulting set is the same.) */
size_t or_enum_free(unsigned n, uint64_t *out, size_t cap) {
    if (n < 1 || n > 8) return 0;
    size_t ncur = 1, cap_cur = 1024;
    uint64_t *cur = malloc(cap_cur * sizeof *cur);
    cur[0] = 1;
    for (unsigned k = 2; k <= n; k++) {
        size_t nn = 0, cap_n = 4096;
        uint64_t *nxt = malloc(cap_n * sizeof *nxt);
        for (size_t i = 0; i < ncur; i++) {
            uint64_t s = cur[i] << 9; /* room above and left (size < 8) */
            for (unsigned b = 0; b < 64; b++) {
                if (!((s >> b) & 1u)) continue;
                int r = (int)(b / 8), c = (int)(b % 8);
                static const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
                for (int d = 0; d < 4; d++) {
                    int rr = r + dr[d], cc = c + dc[d];
                    if (rr < 0 || rr > 7 || cc < 0 || cc > 7) continue;
                    uint64_t bit = 1ull << (rr * 8 + cc);
                    if (s & bit) continue;
                    if (nn == cap_n) {
                        cap_n *= 2;
                        nxt = realloc(nxt, cap_n * sizeof *nxt);
                    }
                    nxt[nn++] = or_canonical(s | bit, 0xffu);
                }
            }
        }
        qsort(nxt, nn, sizeof *nxt, cmp_u64);
        size_t u = 0;
        for (size_t i = 0; i < nn; i++)
            if (!u || nxt[u - 1] != nxt[i]) nxt[u++] = nxt[i];
        free(cur);
        cur = nxt;
        ncur = u;
        (void)cap_cur;
    }
    for (size_t i = 0; i < ncur && i < cap; i++) out[i] = cur[i];
    free(cur);
    return ncur;
}

/* Piece.cpp:12-22: the transforms enabled by sym, in order, each marked
 * disabled when it repeats an earlier one.  Only enabled ones are written;
 * trs[] gets the index within the sym-filtered list, which is what
 * Piece::cover hands to make_step. */
unsigned or_piece_orients(uint64_t shape, unsigned sym, uint64_t *normals,
                          unsigned *trs) {
    uint64_t all[8];
    unsigned nall = 0, nen = 0;
    for (unsigned t = 0; t < 8; t++, sym >>= 1) {
        if (!(sym & 1u)) continue;
        uint64_t sh = or_transform(shape, t);
        int dup = 0;
        for (unsigned j = 0; j < nall; j++)
            if (all[j] == sh) dup = 1;
        all[nall] = sh;
        if (!dup) {
            normals[nen] = sh;
            trs[nen] = nall;
            nen++;
        }
        nall++;
    }
    return nen;
}

void or_step_matrix(unsigned trs, int *m) {
    static const int T[8][4] = {
        {+1, 0, 0, +1}, {-1, 0, 0, +1}, {+1, 0, 0, -1}, {-1, 0, 0, -1},
        {0, +1, +1, 0}, {0, -1, +1, 0}, {0, +1, -1, 0}, {0, -1, -1, 0},
    };
    for (int i = 0; i < 4; i++) m[i] = T[trs & 7][i];
}

/* ------------------------------------------------------------------ */
/* backtracker                                                          */
/* ------------------------------------------------------------------ */

#define OR_MAXP 64

typedef struct {
    uint32_t n, H;
    uint32_t count[OR_MAXP], size[OR_MAXP], used[OR_MAXP];
    unsigned nor[OR_MAXP];
    uint64_t normal[OR_MAXP][8];
    unsigned trs[OR_MAXP][8];
    int maxX[OR_MAXP][8], maxY[OR_MAXP][8];
    uint8_t open[64 + 16];
    uint32_t open_cnt;
    uint64_t max_tiles;
    /* results */
    uint64_t nsol;
    int single, stop;
    uint8_t *out;
    size_t cap, len;
    uint8_t hist[256][4];
    unsigned nhist;
    /* node budget (0 = none) */
    uint64_t nodes, max_nodes;
    int aborted;
} ctx_t;

static uint64_t get_win(const ctx_t *c, unsigned row) {
    uint64_t w;
    memcpy(&w, c->open + row, 8);
    return w;
}
static void put_win(ctx_t *c, unsigned row, uint64_t w) {
    memcpy(c->open + row, &w, 8);
}

static unsigned width8(uint64_t s) {
    unsigned w = 0;
    for (unsigned col = 0; col < 8; col++)
        if (s & (0x0101010101010101ull << col)) w = col + 1;
    return w;
}
static unsigned height8(uint64_t s) {
    unsigned h = 0;
    for (unsigned row = 0; row < 8; row++)
        if (s & (0xffull << (row * 8))) h = row + 1;
    return h;
}

/* Piece.cpp:57-63 */
static uint32_t min_tiles(const ctx_t *c) {
    uint32_t m = UINT32_MAX;
    for (uint32_t i = 0; i < c->n; i++)
        if (c->used[i] < c->count[i] && c->size[i] < m) m = c->size[i];
    return m;
}

static void emit_solution(ctx_t *c) {
    size_t need = 1 + 4u * c->nhist;
    if (c->out && c->len + need <= c->cap) {
        c->out[c->len] = (uint8_t)c->nhist;
        memcpy(c->out + c->len + 1, c->hist, 4u * c->nhist);
    }
    c->len += need;
}

/* one DFS level of solve / solve_count (Piece.cpp:66-139); returns nonzero
 * when a single-mode search has found its solution. */
static int dfs(ctx_t *c, unsigned first_row) {
    if (c->max_nodes && ++c->nodes > c->max_nodes) {
        c->aborted = 1;
        return 1;
    }
    if (c->open_cnt > c->max_tiles) return 0;
    if (!c->open_cnt) {
        c->nsol++;
        if (c->out) emit_solution(c);
        return c->single;
    }
    if (c->open_cnt < min_tiles(c)) return 0;
    unsigned tgtY = first_row;
    while (!c->open[tgtY]) tgtY++;
    unsigned tgtX = (unsigned)__builtin_ctz(c->open[tgtY]);
    for (uint32_t id = 0; id < c->n; id++) {
        if (c->used[id] == c->count[id]) continue;
        c->used[id]++;
        c->max_tiles -= c->size[id];
        for (unsigned o = 0; o < c->nor[id]; o++) {
            uint64_t nm = c->normal[id][o];
            int maxX = c->maxX[id][o], maxY = c->maxY[id][o];
            for (uint64_t bits = nm; bits; bits &= bits - 1) { /* Piece.inl:11 */
                unsigned b = (unsigned)__builtin_ctzll(bits);
                int bitY = (int)(b / 8), bitX = (int)(b % 8);
                if (bitX > (int)tgtX || bitX + maxX < (int)tgtX) continue;
                if (bitY > (int)tgtY || bitY + maxY < (int)tgtY) continue;
                unsigned x = tgtX - (unsigned)bitX, y = tgtY - (unsigned)bitY;
                uint64_t pl = nm << x; /* rows y .. y+7 */
                uint64_t w = get_win(c, y);
                if (pl & ~w) continue;            /* open_tiles >= placed */
                put_win(c, y, w & ~pl);
                unsigned pc = (unsigned)__builtin_popcountll(pl);
                c->open_cnt -= pc;
                c->hist[c->nhist][0] = (uint8_t)id;
                c->hist[c->nhist][1] = (uint8_t)c->trs[id][o];
                c->hist[c->nhist][2] = (uint8_t)y;
                c->hist[c->nhist][3] = (uint8_t)x;
                c->nhist++;
                int f = dfs(c, tgtY);
                c->nhist--;
                c->open_cnt += pc;
                put_win(c, y, get_win(c, y) | pl);
                if (f) return 1;
            }
        }
        c->max_tiles += c->size[id];
        c->used[id]--;
    }
    return 0;
}

/* Placements a piece's `enabled` mask switches on (Piece::Placement::enabled,
 * Piece.hpp:16-20, read by Piece::cover, Piece.inl:8-10): bit p = the p-th
 * transform the sym mask admits, duplicates included when their bit is set. */
static unsigned orients_enabled(uint64_t shape, unsigned sym, unsigned en, uint64_t *normals,
                                unsigned *trs) {
    unsigned nall = 0, k = 0;
    for (unsigned t = 0; t < 8; t++, sym >>= 1) {
        if (!(sym & 1u)) continue;
        if ((en >> nall) & 1u) {
            normals[k] = or_transform(shape, t);
            trs[k] = nall;
            k++;
        }
        nall++;
    }
    return k;
}

static int setup_ex(ctx_t *c, const uint64_t *shapes, const uint32_t *counts,
                    const uint8_t *enabled, uint32_t n, uint32_t sym, const uint8_t *rows,
                    uint32_t H) {
    if (n > OR_MAXP || H > 64) return -1;
    memset(c, 0, sizeof *c);
    c->n = n;
    c->H = H;
    for (uint32_t i = 0; i < n; i++) {
        c->count[i] = counts[i];
        c->size[i] = (uint32_t)__builtin_popcountll(shapes[i]);
        c->nor[i] = enabled ? orients_enabled(shapes[i], sym, enabled[i], c->normal[i], c->trs[i])
                            : or_piece_orients(shapes[i], sym, c->normal[i], c->trs[i]);
        for (unsigned o = 0; o < c->nor[i]; o++) {
            /* p.max = (bottom(), right()) of the normalised shape: the
             * largest translation that stays on the board. */
            c->maxX[i][o] = 8 - (int)width8(c->normal[i][o]);
            c->maxY[i][o] = (int)H - (int)height8(c->normal[i][o]);
        }
        c->max_tiles += (uint64_t)c->size[i] * counts[i];
    }
    for (uint32_t r = 0; r < H; r++) {
        c->open[r] = rows[r];
        c->open_cnt += (uint32_t)__builtin_popcount(rows[r]);
    }
    return 0;
}

static int setup(ctx_t *c, const uint64_t *shapes, const uint32_t *counts,
                 uint32_t n, uint32_t sym, const uint8_t *rows, uint32_t H) {
    return setup_ex(c, shapes, counts, NULL, n, sym, rows, H);
}

uint64_t or_solve_count(const uint64_t *shapes, const uint32_t *counts,
                        uint32_t n, uint32_t sym, const uint8_t *rows,
                        uint32_t H) {
    ctx_t *c = malloc(sizeof *c);
    uint64_t r = 0;
    if (!setup(c, shapes, counts, n, sym, rows, H)) {
        dfs(c, 0);
        r = c->nsol;
    }
    free(c);
    return r;
}

uint64_t or_solve_count_budget(const uint64_t *shapes, const uint32_t *counts,
                               uint32_t n, uint32_t sym, const uint8_t *rows,
                               uint32_t H, uint64_t max_nodes, int *complete) {
    ctx_t *c = malloc(sizeof *c);
    uint64_t r = 0;
    *complete = 0;
    if (!setup(c, shapes, counts, n, sym, rows, H)) {
        c->max_nodes = max_nodes;
        dfs(c, 0);
        r = c->nsol;
        *complete = !c->aborted;
    }
    free(c);
    return r;
}

uint64_t or_solve(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                  uint32_t sym, const uint8_t *rows, uint32_t H, int single,
                  uint8_t *out, size_t cap, size_t *out_len) {
    return or_solve_ex(shapes, counts, NULL, n, sym, rows, H, single, out, cap, out_len);
}

uint64_t or_solve_ex(const uint64_t *shapes, const uint32_t *counts, const uint8_t *enabled,
                     uint32_t n, uint32_t sym, const uint8_t *rows, uint32_t H, int single,
                     uint8_t *out, size_t cap, size_t *out_len) {
    ctx_t *c = malloc(sizeof *c);
    uint64_t r = 0;
    if (!setup_ex(c, shapes, counts, enabled, n, sym, rows, H)) {
        c->single = single != 0;
        c->out = out ? out : (uint8_t *)c; /* non-null: record lengths */
        c->cap = out ? cap : 0;
        dfs(c, 0);
        r = c->nsol;
        if (out_len) *out_len = c->len;
    }
    free(c);
    return r;
}

/* ------------------------------------------------------------------ */
/* record codec (src/record.cuh:5-71)                                   */
/* ------------------------------------------------------------------ */

/* CUDA __byte_perm (PTX prmt, default mode): byte n of the result is byte
 * s[4n+2:4n] of {y:x} (x = bytes 0..3, y = bytes 4..7); selector bit 4n+3
 * replaces it by its sign bit replicated. */
static uint32_t byte_perm(uint32_t x, uint32_t y, uint32_t s) {
    uint64_t v = ((uint64_t)y << 32) | x;
    uint32_t r = 0;
    for (int n = 0; n < 4; n++) {
        uint32_t sel = (s >> (4 * n)) & 0xfu;
        uint32_t b = (uint32_t)(v >> (8 * (sel & 7u))) & 0xffu;
        if (sel & 8u) b = (b & 0x80u) ? 0xffu : 0u;
        r |= b << (8 * n);
    }
    return r;
}

static void ids_to_ex(const uint8_t *nm, uint32_t *ex) {
    for (int w = 0; w < 4; w++)
        ex[w] = (uint32_t)nm[4 * w] | (uint32_t)nm[4 * w + 1] << 8 |
                (uint32_t)nm[4 * w + 2] << 16 | (uint32_t)nm[4 * w + 3] << 24;
}

void or_parse_R(unsigned H, const uint32_t *r, uint8_t ea, uint64_t *empty_area,
                uint32_t *nm_cnt, uint8_t *nm) {
    const uint32_t xaL = r[0], xaH = r[1], ex0 = r[2], ex1 = r[3], ex2 = r[4];
    uint64_t e = 0;
    uint32_t cnt = 0, ex[4] = {~0u, ~0u, ~0u, ~0u};
    if (H == 8) {
        e = ((uint64_t)xaL | (uint64_t)xaH << 32) << 8 | ea;
        *empty_area = e;
        *nm_cnt = 0;
        memset(nm, 0xff, 16);
        return;
    }
    if (H == 7) {
        e = xaL | (uint64_t)(xaH & 0x0000ffffu) << 32;
        cnt = xaH >> 24;
    }
    if (H == 6) {
        e = xaL | (uint64_t)(xaH >> 16 & 0xffu) << 32;
        cnt = xaH >> 24;
        ex[3] = xaH | 0xffff0000u;
    }
    if (H == 5) {
        e = xaL;
        cnt = xaH >> 24;
        ex[3] = xaH | 0xff000000u;
    }
    if (H <= 4) {
        e = xaL & 0x00ffffffu;
        cnt = xaL >> 24;
        ex[3] = xaH;
    }
    if (H <= 6) ex[2] = ex2;
    ex[1] = ex1;
    ex[0] = ex0;
    *empty_area = e << 8 | ea;
    *nm_cnt = cnt;
    for (int w = 0; w < 4; w++)
        for (int b = 0; b < 4; b++) nm[4 * w + b] = (uint8_t)(ex[w] >> (8 * b));
}

void or_assemble_R(unsigned H, uint64_t empty_area, uint32_t nm_cnt, const uint8_t *nm,
                   uint32_t *r, uint8_t *ea) {
    uint32_t ex[4];
    ids_to_ex(nm, ex);
    uint32_t xaL = 0, xaH = 0, ex2 = ~0u;  /* ex2 unset at H == 7 (record.cuh:47) */
    *ea = (uint8_t)(empty_area >> 8 & 0xffu);
    if (H == 7) {
        xaL = (uint32_t)(empty_area >> 16);
        xaH = (uint32_t)(empty_area >> 48) | nm_cnt << 24;
    }
    if (H == 6) {
        xaL = (uint32_t)(empty_area >> 16);
        xaH = byte_perm(byte_perm(ex[3], nm_cnt, 0x4f10u), (uint32_t)(empty_area >> 32), 0x3610u);
    }
    if (H == 5) {
        xaL = (uint32_t)(empty_area >> 16);
        xaH = byte_perm(ex[3], nm_cnt, 0x4210u);
    }
    if (H <= 4) {
        xaL = byte_perm((uint32_t)(empty_area >> 16), nm_cnt, 0x4210u);
        xaH = ex[3];
    }
    if (H <= 6) ex2 = ex[2];
    r[0] = xaL;
    r[1] = xaH;
    r[2] = ex[0];
    r[3] = ex[1];
    r[4] = ex2;
}
