/*
 * d4.h -- C ABI of the MI355X (gfx950) row-by-row polyomino tiling search.
 *
 * Drop-in counterpart of the reference's GPU row driver
 *   struct CudaSearcher                       src/searcher_cuda.h:10-38
 * plus the set-up it needs (piece library, placement tables, device upload:
 *   compute_frow_on_cpu / transfer_frow_to_gpu / prepare_kernels,
 *   src/frow.h:10-12, src/kernel.h:80) and the terminal stage the reference
 * left unfinished (src/searcher.cpp:36-65).
 *
 * Conventions: every call returns D4_OK (0) or a negative D4_E* code and
 * never throws; d4_last_error() gives the calling thread's last message.
 * Handles are owned by the caller and freed with the matching _destroy;
 * device memory is owned by the library.  A handle is used by one thread at
 * a time ("NOT thread-safe", src/searcher.hpp:15-16); distinct handles may be
 * used from distinct threads.
 *
 * Board encoding: row-major, 8 bits per row (bit c of rows[r] = cell (r, c)
 * is open / must be covered), width <= 8, height <= 64.  This is the
 * reference's Shape<8> layout (bit = row*8 + col, src/Shape.hpp:20-23)
 * extended downwards.
 */
#ifndef D4_H
#define D4_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    D4_OK = 0,
    D4_EINVAL = -1,   /* bad argument / unsupported shape of problem      */
    D4_ENOMEM = -2,   /* device or host allocation failed                 */
    D4_EDEVICE = -3,  /* HIP runtime error (no device, launch failure ...) */
    D4_ESTATE = -4,   /* call not valid in the handle's current state     */
    D4_ERANGE = -5,   /* a capacity or counter limit was exceeded         */
};

enum {
    D4_MODE_COUNT = 0, /* count tilings: frontier states carry u64 counts  */
    D4_MODE_ENUM = 1,  /* also keep per-row back-pointers, so the full
                          solution set and the first solution in the CPU
                          searcher's DFS order can be reconstructed        */
};

enum {
    D4_FLAG_NO_PRUNE = 1, /* disable region-connectivity pruning           */
    D4_FLAG_HOST_TABLES = 2, /* multi-library batch: build the per-library
                                placement tables on the host (default: on
                                the device)                                */
};

typedef struct d4_library d4_library;
typedef struct d4_searcher d4_searcher;
typedef struct d4_batch d4_batch;

typedef struct d4_stats {
    uint64_t rows_done;        /* row steps executed                        */
    uint64_t records_extended; /* parents expanded, summed over rows        */
    uint64_t children;         /* children emitted (before dedup)           */
    uint64_t unique;           /* unique states kept (after dedup)          */
    uint64_t pruned;           /* children rejected by connectivity pruning */
    uint64_t peak_frontier;    /* largest frontier (records)                */
    uint64_t retries;          /* row steps re-run after a table overflow   */
    double   gpu_ms;           /* device time of the row steps              */
    double   extend_ms;        /* device time inside the extend kernel      */
    uint64_t extend_launches;  /* extend launches (retries included)        */
    uint64_t extend_bytes;     /* algorithmic HBM bytes of those launches:
                                  16 per parent read + 16 per child merged  */
    double   compact_ms;       /* device time inside the compaction kernel  */
    uint64_t compact_launches;
    uint64_t compact_bytes;    /* 8 per table slot scanned + 24 per unique
                                  state gathered and written                */
    uint64_t count_passes;     /* search-only passes sizing the dedup table */
    uint64_t splits;           /* frontier halves parked to stay in the memory cap */
    uint64_t donated;          /* work items handed between extend launches */
    uint64_t rehashes;         /* dedup tables grown after children spilled */
    uint64_t sink_growths;     /* terminal-state sinks grown (the step reruns) */
    uint64_t host_syncs;       /* host round trips (waits on the device) inside
                                  the row steps                             */
    double   dedup_ms;         /* device time of the separate dedup kernel
                                  (level kernels; 0 where dedup is fused
                                  into the extend kernel)                   */
    uint64_t dedup_launches;
    uint64_t dedup_bytes;      /* 16 per child read + 16 per new state      */
    uint64_t chunk_reruns;     /* level-kernel chunks rerun with larger buffers */
    uint64_t donation_full;    /* donations / spills refused by a full buffer */
    uint64_t fused_steps;      /* row steps queued as one chain with a single
                                  host round trip (level kernels, count mode) */
    uint64_t fused_fallbacks;  /* of those, steps rerun on the general path
                                  (a buffer overflowed or children spilled) */
    uint64_t graph_replays;    /* fused row steps replayed from a HIP graph  */
    uint64_t host_parked_peak; /* most bytes of parked parts in host memory  */
    uint64_t streamed_steps;   /* row steps run in hash partitions (too large
                                  for one dedup table under the cap)        */
    uint64_t stream_passes;    /* partitions of those steps                 */
    uint64_t part_steps;       /* row steps deduped by hash buckets merged in
                                  LDS                                       */
    uint64_t part_fallbacks;   /* of those, steps rerun with the device-memory
                                  table (more children than the buckets' buffers
                                  were sized for, a coarse bucket past its
                                  tiles, or a hash sub-range with more unique
                                  states than an LDS table holds)           */
    uint64_t part_remerges;    /* bucket merges rerun alone into a larger next
                                  frontier (more unique states than estimated) */
} d4_stats;

/* Thread-local text of the last error. */
const char *d4_last_error(void);

/* Number of visible HIP devices (0 when none). */
int d4_device_count(int *n);

/* Wait for all work on the current device (hipDeviceSynchronize). */
int d4_device_synchronize(void);

/* Return the device memory the library's caching allocator holds for reuse
 * on the current device to the runtime (after the device's pending work);
 * bytes (may be NULL) receives how much.  Handles keep their own buffers.
 * Useful between workloads of very different size: a later allocation that
 * finds no cached block of its size would otherwise release the whole cache
 * at once, a multi-second pause after a search of hundreds of GB.  No
 * counterpart in the reference (its buffers come from cudaMalloc per row). */
int d4_device_release_cached(uint64_t *bytes);

/* Measured rate of random 64-bit device-scope atomic adds into 1 GiB of
 * device memory (best of three launches of 134 M adds): the ceiling of a
 * dedup that merges each child into a hash table in HBM with one atomic.
 * No counterpart in the reference (measurement only). */
int d4_device_atomic_rate(double *ops_per_s);

/* Free polyominoes with n cells (1 <= n <= 8), each as its canonical value
 * (minimum over the 8 D4 transforms of the normalised 8-wide encoding),
 * ascending: the reference's known_shapes[n] tables (src/known1.cpp:3-14,
 * generated by src/enum.cpp:28-57).  Writes min(total, cap) values and
 * returns the total.  Host-only; needs no device. */
size_t d4_free_polyominoes(unsigned n, uint64_t *out, size_t cap);

/* ---------------------------------------------------------------------- */
/* piece library and placement tables                                      */
/* ---------------------------------------------------------------------- */

/* Build the per-anchor placement tables for a library of n pieces
 * (shapes[i] in the 8-wide encoding, normalised; counts[i] = multiplicity)
 * under symmetry mask sym (0xff = all D4 transforms, src/Group.hpp:8-19) for
 * boards of `width` columns, and upload them to the current device.
 * Replaces compute_frow_on_cpu + transfer_frow_to_gpu + prepare_kernels
 * (src/frow.cpp:39-166, src/frow.cu:7-35, src/kernel.cu:73-93); the tables
 * are full-row (no L/R split) and carry no static island pruning.
 * Limits: n <= 32, <= 64 unique (piece, transform) pairs, piece height <= 6,
 * sum of per-piece counter widths + state bits <= 63. */
int d4_library_create(const uint64_t *shapes, const uint32_t *counts,
                      uint32_t n, uint32_t sym, uint32_t width,
                      d4_library **out);
/* d4_library_create with each piece's placements switched on by its own
 * mask, as the UI does through Piece::Placement::enabled (src/Piece.hpp:16-20,
 * bound in src/wasm.cpp:37-47; read by Piece::cover, src/Piece.inl:8-10):
 * bit p of enabled[i] = the p-th transform sym admits (the placement's
 * trs_id); a duplicate shape is a placement of its own when its bit is set.
 * enabled == NULL is d4_library_create (duplicates off). */
int d4_library_create_ex(const uint64_t *shapes, const uint32_t *counts, uint32_t n,
                         uint32_t sym, const uint8_t *enabled, uint32_t width,
                         d4_library **out);
void d4_library_destroy(d4_library *lib);
/* n_cand = unique (piece, transform) pairs, max_height = tallest piece. */
int d4_library_info(const d4_library *lib, uint32_t *n_cand,
                    uint32_t *max_height);

/* Launch shape of the extend kernel: fit_rows = window rows covered by the
 * LDS fit table (1..max_height; the chosen placement is always tested in
 * full), threads = workgroup size (256, 512 or 1024), refill = idle lanes
 * of a wave that trigger a refill round (1..64).  Not to be changed while a
 * search on this library is running.  Replaces the reference's KParams
 * choice (src/kernel.h:44-67). */
int d4_library_set_launch(d4_library *lib, uint32_t fit_rows, uint32_t threads,
                          uint32_t refill);
/* Load balance of the extend kernel: a lane that has run `steps` DFS steps
 * on one work item hands the rest of its search tree to a follow-up launch
 * (0: never).  d4_library_tune sets it too.  (KSizing::optimize has no
 * counterpart: the reference's tiles are fixed per launch, src/kernel.cu.) */
int d4_library_set_budget(d4_library *lib, uint32_t steps);
int d4_library_get_budget(const d4_library *lib, uint32_t *steps);
/* Kernel of count-mode row steps (searchers, batches, the server's
 * all-solutions frontier, bounded runs).
 *   D4_KERNEL_LEVEL (the default): level-synchronous -- one launch per
 *     placement level over every partial row of the frontier, (row,
 *     candidate) pairs spread one per lane, partial rows and children
 *     written densely to device memory, then a separate dedup kernel
 *     (prune + table merge).  Needs fit tables of at most 4 rows.
 *   D4_KERNEL_EXPAND: partial-row snapshots on a per-wave LDS stack, one
 *     candidate per lane per round; a wave's snapshots move to a follow-up
 *     launch once the launch's records are gone and the wave has run
 *     tail_rounds more rounds (0: never).
 *   D4_KERNEL_DFS: one depth-first search per lane with the donation budget
 *     of d4_library_set_budget.
 * All replace the reference's tiled_row_search (src/kernel.cuh:79-160);
 * counts are identical.  Enum mode (solutions) always uses the depth-first
 * kernel. */
enum { D4_KERNEL_DFS = 0, D4_KERNEL_EXPAND = 1, D4_KERNEL_LEVEL = 2 };
int d4_library_set_kernel(d4_library *lib, uint32_t kernel, uint32_t tail_rounds);
int d4_library_get_kernel(const d4_library *lib, uint32_t *kernel, uint32_t *tail_rounds);
/* How a count-mode row step on the level kernels is issued (no counterpart in
 * the reference, whose launch loop is host-driven per row: src/device.cu:39-162):
 *   D4_STEP_GENERAL: host round trips between the search, the table checks and
 *     the compaction (about four per row step);
 *   D4_STEP_FUSED (the default): the whole row step -- table fill, counter
 *     resets, terminal pass, level launches, dedup, compaction -- queued as one
 *     chain with one host round trip; a step whose buffers overflow or whose
 *     children spill is rerun on the general path;
 *   D4_STEP_GRAPH: the fused chain captured once into a HIP graph and replayed
 *     row after row, its sizes and buffer choice read from device memory. */
enum { D4_STEP_GENERAL = 0, D4_STEP_FUSED = 1, D4_STEP_GRAPH = 2 };
int d4_library_set_step_mode(d4_library *lib, uint32_t mode);
int d4_library_get_step_mode(const d4_library *lib, uint32_t *mode);
/* Dedup of count-mode level row steps, fused or replayed (replaces the
 * reference's host Sorter, src/sorter.cpp:156-214, whose set semantics it
 * keeps: equal states merge, their tiling counts summed):
 *   D4_DEDUP_TABLE: every child merged into one open-addressing table in
 *     device memory (a coherent load, an atomic add, a CAS for a new state),
 *     then the table compacted into the next frontier;
 *   D4_DEDUP_PART (the default): children pruned and the survivors moved by
 *     their key's hash into buckets (a histogram, a scan and a scatter, no
 *     device atomic per child), each bucket merged in one workgroup's LDS
 *     hash table and its unique states written to the next frontier; rows
 *     with more than 2^30 children use the table.
 * Counts are identical; only the order of the next frontier differs. */
enum { D4_DEDUP_TABLE = 0, D4_DEDUP_PART = 1 };
int d4_library_set_dedup(d4_library *lib, uint32_t mode);
int d4_library_get_dedup(const d4_library *lib, uint32_t *mode);
/* Items per donation buffer of the depth-first and snapshot kernels (default
 * 1 Mi; at least 64).  A lane or wave that finds the buffer full keeps its
 * work (counted in d4_stats.donation_full); small values exercise that path. */
int d4_library_set_donation_capacity(d4_library *lib, uint32_t items);
int d4_library_get_launch(const d4_library *lib, uint32_t *fit_rows,
                          uint32_t *threads, uint32_t *refill);
/* Measured autotuner, replacing the analytic cost model of
 * KSizing::optimize / KParams::fom (src/kernel.cu:95-236): solves the given
 * boards (count mode, as d4_batch_*) once per candidate launch shape, keeps
 * the shape with the least extend-kernel time and sets it on the library.
 * Outputs the chosen shape and its extend time in ms (any may be NULL). */
int d4_library_tune(d4_library *lib, uint32_t n_inst, const uint8_t *boards,
                    uint32_t stride, const uint32_t *heights,
                    uint32_t *fit_rows, uint32_t *threads, uint32_t *refill,
                    double *best_ms);

/* ---------------------------------------------------------------------- */
/* single-instance searcher (CudaSearcher)                                 */
/* ---------------------------------------------------------------------- */

/* CudaSearcher::CudaSearcher(uint64_t empty_area), src/searcher_cuda.cu:12-21:
 * an 8x8 board given as one u64 (bit = row*8+col set = open); height =
 * ceil(bit_width(empty_area) / 8) as in the reference. */
int d4_searcher_create_ea(const d4_library *lib, uint64_t empty_area,
                          uint32_t mode, uint32_t flags, d4_searcher **out);
/* General board: `height` rows of 8-bit masks. */
int d4_searcher_create(const d4_library *lib, const uint8_t *rows,
                       uint32_t height, uint32_t mode, uint32_t flags,
                       d4_searcher **out);
/* ~CudaSearcher, src/searcher_cuda.cu:23-30. */
void d4_searcher_destroy(d4_searcher *s);
/* get_height(), src/searcher_cuda.h:14: rows still to be filled. */
uint32_t d4_searcher_height(const d4_searcher *s);
/* size(), src/searcher_cuda.h:16-21: records in the current frontier. */
uint64_t d4_searcher_size(const d4_searcher *s);
/* next_size(), src/searcher_cuda.h:22-27, src/searcher_cuda.cu:84-90: an
 * upper bound on the children the next step can emit (sum over records of
 * the number of ways to fill the record's current row ignoring lower rows and
 * piece counts). */
uint64_t d4_searcher_next_size(const d4_searcher *s);
/* size(unsigned i), src/searcher_cuda.h:16-21: records of the current
 * frontier whose current row has open cells `ea` (the reference's WL bucket
 * `pos`, 0..255; bit c = column c open).  0 for ea > 255 or on error. */
uint64_t d4_searcher_size_bucket(const d4_searcher *s, unsigned ea);
/* next_size(unsigned i), src/searcher_cuda.cu:84-90: the bound of
 * d4_searcher_next_size restricted to bucket `ea`.  Summed over the 256
 * buckets both equal d4_searcher_size / d4_searcher_next_size. */
uint64_t d4_searcher_next_size_bucket(const d4_searcher *s, unsigned ea);
/* The current frontier's states as the engine keeps them: keys[k] = the
 * occupancy by placed pieces of the current row and the rows below it (8
 * bits per row, max_height - 1 rows, d4_library_info) followed by the
 * library's packed use counters (d4_library_counters), counts[k] = the
 * state's tiling count; states first .. first + cap - 1 in the frontier's
 * order (dedup-table order, which does not follow the states' values).
 * *n = frontier size.  keys / counts may be NULL. */
int d4_searcher_frontier(const d4_searcher *s, uint64_t first, uint64_t cap, uint64_t *keys,
                         uint64_t *counts, uint64_t *n);
/* Keep up to cap states (keys and counts as d4_searcher_frontier) of the
 * frontier the search reaches at `height` rows left, taken as d4_searcher_run
 * (bounded runs included) or d4_searcher_step pass it; 0 turns it off.  With
 * a split frontier the states come from the parts that reach the height
 * first.  d4_searcher_sample returns them (*n = states kept). */
int d4_searcher_set_sample(d4_searcher *s, uint32_t height, uint64_t cap);
int d4_searcher_sample(const d4_searcher *s, uint64_t *keys, uint64_t *counts, uint64_t cap,
                       uint64_t *n);
/* search_GPU(), src/searcher_cuda.cu:32-82: advance the frontier by one row
 * (extend + prune + dedup + compact on the device). */
int d4_searcher_step(d4_searcher *s);
/* Step until height == 0. */
int d4_searcher_run(d4_searcher *s);
/* Count mode: device bytes the frontier and its dedup table may take in
 * d4_searcher_run (default: 80% of the free device memory at creation).  A
 * row that would exceed it parks half of the frontier and finishes the
 * other half first (depth first), streaming a frontier larger than memory
 * through it; the count is unchanged.  d4_searcher_step is not bounded. */
int d4_searcher_set_memory(d4_searcher *s, uint64_t bytes);
/* Count mode: pinned host bytes the parked frontier parts of
 * d4_searcher_run may take (default 0: parts stay in device memory).  With a
 * budget, a part is copied out to host memory when parked and back when its
 * turn comes, so the frontier of a row can exceed the device memory -- the
 * role of the reference's ring buffer drained to host memory
 * (src/device.cu:32, :209-230); D4_ENOMEM when the budget runs out. */
int d4_searcher_set_host_memory(d4_searcher *s, uint64_t bytes);
/* Allocation-free count: the number of tilings of one board (as
 * d4_searcher_create + d4_searcher_run + d4_searcher_count in count mode),
 * computed without allocating device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Its
 * buffers are slices of one 128 MiB device array per GPU (set up when the
 * first library is created on that GPU and held until the process ends):
 * at most 512 Ki frontier states, a 2 Mi-slot dedup table, 512 Ki partial
 * rows per level and 2 Mi children per chunk; a search that needs more
 * fails with D4_ERANGE.  The library's tables must be on the device
 * (libraries of at most 64 (piece, transform) pairs are uploaded at
 * creation).  Every call on a GPU shares that GPU's array, which keeps the
 * last call's contents (each call clears what must start at zero): calls
 * in one process are serialised by a lock; callers that share the array in
 * other ways race. */
int d4_searcher_count_noalloc(const d4_library *lib, const uint8_t *rows,
                              uint32_t height, uint32_t flags, uint64_t *count);
/* Number of tilings (valid once height == 0): sum of terminal counts. */
int d4_searcher_count(const d4_searcher *s, uint64_t *count);
/* Solutions (D4_MODE_ENUM only, height == 0), in the CPU searcher's DFS
 * order (solve(), src/Piece.cpp:66-102).  Each is written as
 *   u8 nsteps, nsteps x {u8 piece_id, u8 trs_id, u8 y, u8 x}
 * (Step<L>::piece_id, trs_id, y, x, src/Piece.hpp:31-36).  single != 0
 * writes only the first solution (solve(..., single=true)).  *len gets the
 * bytes needed, *nsol the solutions written (or that would be). */
int d4_searcher_solutions(const d4_searcher *s, int single, uint8_t *out,
                          size_t cap, size_t *len, uint64_t *nsol);
int d4_searcher_stats(const d4_searcher *s, d4_stats *st);
/* The current frontier in the reference's record format (src/record.h:5-23):
 * one 24-byte RX per state -- R {xaL, xaH, ex0, ex1, ex2} as parse_R<H>
 * reads it (src/record.cuh:5-41; H = rows left including the current one)
 * and ea = the current row's open cells -- with piece ids = library indices,
 * sorted, 0xff padded.  Ordered by (ea, R).  counts (optional) receives each
 * state's tiling count (the reference keeps none).  Only the reference's
 * domain can be written: 8x8 boards (H <= 8), at most 16 pieces, counts 1,
 * else D4_ERANGE.  *n gets the number of records (call with cap 0 first). */
int d4_searcher_export(const d4_searcher *s, void *rx, size_t cap, size_t *n,
                       uint64_t *counts);

/* ---------------------------------------------------------------------- */
/* batched instances (no reference counterpart: the reference runs one     */
/* board config per CudaSearcher, src/searcher.cpp:68-107)                 */
/* ---------------------------------------------------------------------- */

/* n_inst boards sharing one library; boards[i*stride + r] = row r of board
 * i, heights[i] its row count.  All instances advance row by row in one
 * shared frontier (one launch per kernel per row for all of them). */
int d4_batch_create(const d4_library *lib, uint32_t n_inst,
                    const uint8_t *boards, uint32_t stride,
                    const uint32_t *heights, uint32_t mode, uint32_t flags,
                    d4_batch **out);
void d4_batch_destroy(d4_batch *b);
int d4_batch_run(d4_batch *b);
/* Count mode: memory cap of d4_batch_run, as d4_searcher_set_memory. */
int d4_batch_set_memory(d4_batch *b, uint64_t bytes);
/* Count mode: host budget for parked parts, as d4_searcher_set_host_memory. */
int d4_batch_set_host_memory(d4_batch *b, uint64_t bytes);
/* counts[i] = number of tilings of board i (valid after d4_batch_run). */
int d4_batch_counts(const d4_batch *b, uint64_t *counts);
int d4_batch_stats(const d4_batch *b, d4_stats *st);
/* Solutions of instance i (D4_MODE_ENUM), format as d4_searcher_solutions. */
int d4_batch_solutions(const d4_batch *b, uint32_t i, int single,
                       uint8_t *out, size_t cap, size_t *len, uint64_t *nsol);

/* ---------------------------------------------------------------------- */
/* multi-library batches: every instance has its own piece set and board   */
/* width (no reference counterpart: the reference's library is a process   */
/* global, g_nme / g_sym, src/frow.h:8-9, set once per run of              */
/* SearcherFactory::run, src/searcher.cpp:86-107)                          */
/* ---------------------------------------------------------------------- */

typedef struct d4_mbatch d4_mbatch;

/* n_inst instances; libs[i] = instance i's library (libraries may repeat and
 * may hold up to 128 (piece, transform) pairs; rows are within libs[i]'s
 * width), boards[i*stride + r] = row r, heights[i] = row count.  Count mode
 * (all tilings).  At most 32768 instances. */
/* Bytes of one library's placement tables in a multi-library batch. */
uint64_t d4_mlib_table_bytes(void);
/* Instance inst's library tables as the batch holds them on the device, and
 * the same tables built by the host from a library, for comparing the two
 * builds.  The reference builds its row tables on the host once per run and
 * copies them over (compute_frow_on_cpu, src/frow.cpp:39-166;
 * transfer_frow_to_gpu, src/frow.cu:7); a batch builds one table set per
 * distinct library on the device (D4_FLAG_HOST_TABLES: on the host). */
int d4_mbatch_tables(const d4_mbatch *b, uint32_t inst, void *out, uint64_t bytes);
int d4_library_host_tables(const d4_library *lib, void *out, uint64_t bytes);
int d4_mbatch_create(uint32_t n_inst, const d4_library *const *libs,
                     const uint8_t *boards, uint32_t stride,
                     const uint32_t *heights, uint32_t flags, d4_mbatch **out);
void d4_mbatch_destroy(d4_mbatch *b);
/* Extend workgroup size (256, 512, 1024) and refill threshold (1..64). */
int d4_mbatch_set_launch(d4_mbatch *b, uint32_t threads, uint32_t refill);
/* Measured autotuner: runs a batch of the first `sample` instances (0: 256)
 * with each workgroup size and refill threshold, keeps the fastest on b. */
int d4_mbatch_tune(d4_mbatch *b, uint32_t sample, uint32_t *threads,
                   uint32_t *refill, double *best_ms);
/* DFS steps a lane runs on one work item of a row's main launch before it
 * hands the rest of its search tree to follow-up launches, as
 * d4_library_set_budget (default 4096; 0: never). */
int d4_mbatch_set_budget_steps(d4_mbatch *b, uint32_t steps);
/* Items per donation buffer (default 1 Mi; at least 64), as
 * d4_library_set_donation_capacity. */
int d4_mbatch_set_donation_capacity(d4_mbatch *b, uint32_t items);
/* Kernel of the batch's row steps (replacing tiled_row_search,
 * src/kernel.cuh:79-160, for instances of different libraries):
 *   D4_KERNEL_LEVEL (the default): level-synchronous -- one launch per
 *     placement level over every partial row of the frontier (each row with
 *     its instance's library), children written densely and pruned per
 *     wave, then a separate insert into the two-word dedup table; the
 *     frontier goes through in chunks sized to an eighth of the memory cap.
 *   D4_KERNEL_DFS: one depth-first search per lane with donation of the
 *     rest of a long search to follow-up launches (d4_mbatch_set_budget_steps,
 *     d4_mbatch_set_donation_capacity; d4_mbatch_set_launch and
 *     d4_mbatch_tune set its launch shape).
 * Counts are identical; D4_KERNEL_EXPAND is refused (D4_EINVAL). */
int d4_mbatch_set_kernel(d4_mbatch *b, uint32_t kernel);
/* Device bytes the frontier, its dedup table and parked frontier parts may
 * take (default: 80% of the free device memory when the batch is created).
 * A row step that would exceed it parks half of the frontier and finishes
 * the other half first; counts are unchanged. */
int d4_mbatch_set_memory(d4_mbatch *b, uint64_t bytes);
/* States one instance may hold in a frontier (0, the default: no limit).
 * An instance that exceeds it leaves the search unfinished: its count is 0
 * and d4_mbatch_instances reports it. */
int d4_mbatch_set_budget(d4_mbatch *b, uint64_t max_states);
int d4_mbatch_run(d4_mbatch *b);
/* counts[i] = number of tilings of instance i (valid after d4_mbatch_run). */
int d4_mbatch_counts(const d4_mbatch *b, uint64_t *counts);
/* Per instance: the largest number of its states in one frontier, and
 * whether its count is complete (1) or it exceeded the budget (0). */
int d4_mbatch_instances(const d4_mbatch *b, uint64_t *peak_states,
                        uint8_t *finished);
int d4_mbatch_stats(const d4_mbatch *b, d4_stats *st);
/* Keep the terminal states (every cell covered) that instances reach, for
 * d4_mbatch_terminals (off by default).  The device buffer starts with
 * room for initial_entries states (0: 4 per instance, at least 4096) and
 * grows as needed (a row step that overflows it is rerun).  Replaces the
 * records the
 * commented-out consumer of src/searcher.cpp:38-65 reads from the
 * CudaSearcher frontier once its height reaches 0. */
int d4_mbatch_set_terminals(d4_mbatch *b, int keep, uint64_t initial_entries);
/* After d4_mbatch_run with terminals kept: one entry per distinct (instance,
 * piece use counters) terminal state of a finished instance, sorted by
 * instance then counters; tilings[k] = number of tilings ending in it.
 * Counters are packed as the library packs them (d4_library_counters).
 * Writes min(total, cap) entries and sets *n = total. */
int d4_mbatch_terminals(const d4_mbatch *b, uint32_t *inst, uint64_t *counters,
                        uint64_t *tilings, size_t cap, size_t *n);
/* Bit offset and width of each piece's use counter in a library's packed
 * counters (n entries each, n = the library's piece count). */
int d4_library_counters(const d4_library *lib, uint32_t *offset, uint32_t *width,
                        uint32_t *n);

/* ---------------------------------------------------------------------- */
/* piece-set names and ledgers: the consumer of terminal states            */
/* (src/naming.hpp:15-71, src/naming.inl:6-72, src/naming.cpp:9-81,        */
/*  src/search.cpp:25-226)                                                  */
/* ---------------------------------------------------------------------- */

/* Canonical polyominoes with n cells (1 <= n <= 8) under the transforms in
 * `sym` (bit t = D4 transform t): each as the minimum over those transforms
 * of the normalised 8-wide encoding, ascending.  sym 0xFF gives the free
 * tables (src/known1.cpp:3-14); sym 0x69 (identity and the three rotations)
 * gives one-sided ones (src/known2.cpp:3-14, generated by
 * src/enum.cpp:28-57).  Writes min(total, cap) and returns the total. */
size_t d4_polyominoes(unsigned n, uint32_t sym, uint64_t *out, size_t cap);

typedef struct d4_naming d4_naming;
#define D4_NO_NAME UINT64_MAX

/* Naming(m, M, n, N, t, arr, sz) (src/naming.hpp:17, src/naming.cpp:35-81):
 * the piece sets that take between min_n and max_n pieces of sizes min_m ..
 * max_m, at most one of each, arr_sizes[m] pieces of size m (arr_sizes has
 * max_m + 1 entries; index < min_m unused), covering tgt cells in all.
 * Names are ranks: partitions (pieces per size) in lexicographic order of
 * their per-size counts, then per size, from min_m (most significant) to
 * max_m, the rank of the chosen indices among k-subsets in lexicographic
 * order.  D4_ERANGE when a count exceeds 2^64. */
int d4_naming_create(uint64_t min_m, uint64_t max_m, uint64_t min_n, uint64_t max_n,
                     uint64_t tgt, const uint64_t *arr_sizes, d4_naming **out);
void d4_naming_destroy(d4_naming *nm);
/* Naming::size / size_pieces (src/naming.hpp:30-39). */
uint64_t d4_naming_size(const d4_naming *nm);
uint64_t d4_naming_size_pieces(const d4_naming *nm);
/* Naming::name_piece (src/naming.hpp:20-24): the id of piece i of size m;
 * D4_NO_NAME when out of range. */
uint64_t d4_naming_name_piece(const d4_naming *nm, uint64_t m, uint64_t i);
/* Naming::resolve_piece (src/naming.cpp:25-33); D4_ERANGE (with m = i = 0)
 * for an id >= size_pieces. */
int d4_naming_resolve_piece(const d4_naming *nm, uint64_t id, uint64_t *m, uint64_t *i);
/* Naming::name (src/naming.inl:6-29): used[id] != 0 for each piece id in the
 * set (size_pieces entries); *name = D4_NO_NAME when the set is not one of
 * the named ones (the reference's empty optional). */
int d4_naming_name(const d4_naming *nm, const uint8_t *used, uint64_t *name);
/* Naming::resolve (src/naming.inl:61-72): used[id] = 1 for the pieces of
 * the named set, 0 otherwise; D4_ERANGE for name >= size (the reference's
 * false). */
int d4_naming_resolve(const d4_naming *nm, uint64_t name, uint8_t *used);
/* Names of n terminal states of a search over a library whose piece p is
 * naming piece id p (a library of size_pieces pieces): counters[k] is
 * decoded with the library's counter layout; a state that uses a piece
 * more than once, or a set outside the naming, gets D4_NO_NAME (the
 * reference prints "Naming rejected" for those, src/searcher.cpp:58-60). */
int d4_naming_name_counters(const d4_naming *nm, const d4_library *lib,
                            const uint64_t *counters, size_t n, uint64_t *names);

typedef struct d4_ledger d4_ledger;
/* Ledger kinds, the reference's S=v|ac|fb|fB (src/search.cpp:297-315). */
enum {
    D4_LEDGER_VERIFY = 0, /* Verify: counts work only (search.cpp:25-35)    */
    D4_LEDGER_COUNT = 1,  /* AtomicCount: u64 per name (search.cpp:37-54)   */
    D4_LEDGER_BIT = 2,    /* FileBit: file, one bit per (name, config)      */
    D4_LEDGER_BYTE = 3    /* FileByte: file, one byte per (name, config)    */
};

/* A ledger over n_configs board configurations and n_names piece sets.
 * The two file kinds map `path` shared (FileSearcherBase,
 * src/search.cpp:56-174) with the reference's layout: n_configs 32-byte
 * descriptors {shape, count, solved (ms, 0 = unsolved), 0}, then n_names
 * u64 per-name config counts, then the ledger: n_names rows of n_configs
 * bytes (FileByte, 0xFF = set) or ceil(n_configs / 8) bytes (FileBit, bit
 * c % 8 of byte c / 8).  An empty file is extended to that size; a longer
 * one is refused ("file too long"), a shorter non-empty one too ("file too
 * short"); D4_EINVAL. */
int d4_ledger_open(int kind, const char *path, uint64_t n_configs, uint64_t n_names,
                   d4_ledger **out);
/* Unmaps and closes (the mapping is shared: the file holds every change). */
void d4_ledger_close(d4_ledger *L);
/* SearcherFactory::should_run (src/searcher.hpp:70, search.cpp:156-159):
 * the file kinds record the config's shape and run it only when unsolved. */
int d4_ledger_should_run(d4_ledger *L, uint64_t config, uint64_t shape, int *run);
/* Searcher::log (src/search.cpp:28-31, 44-50, 182-191, 207-220) for one
 * (config, name) pair: *fresh = 1 when it was not logged for the config
 * before (then the work counter advances), else 0.  Names D4_NO_NAME are
 * skipped (*fresh = 0). */
int d4_ledger_log(d4_ledger *L, uint64_t config, uint64_t name, int *fresh);
/* log for n pairs; fresh[k] as d4_ledger_log (may be NULL). */
int d4_ledger_log_many(d4_ledger *L, const uint64_t *configs, const uint64_t *names,
                       size_t n, uint8_t *fresh);
/* SearcherFactory::after_run (search.cpp:161-165, searcher.cpp:110-115):
 * the file kinds store the config's count (log calls that returned 1) and
 * the milliseconds it took (at least 1, so that it reads as solved). */
int d4_ledger_after_run(d4_ledger *L, uint64_t config, uint64_t shape, uint64_t count,
                        uint64_t ms);
/* Work counter (log calls that returned 1) and configs finished. */
int d4_ledger_progress(const d4_ledger *L, uint64_t *work, uint64_t *configs_done);
/* Per-name totals: AtomicCount's ledger or the file kinds' name header
 * (configs each name was logged for); Verify has none (D4_ESTATE).  Writes
 * min(n_names, cap) values. */
int d4_ledger_names(const d4_ledger *L, uint64_t *out, size_t cap);

/* ---------------------------------------------------------------------- */
/* serving: continuous batching of requests (no reference counterpart)     */
/* ---------------------------------------------------------------------- */

typedef struct d4_server d4_server;

/* A request loop over one library.  `slots` requests of each kind can be in
 * flight at once ("solve all" requests run in a count-mode frontier, "solve
 * one" requests in an enum-mode frontier); boards up to max_height rows.
 * Requests are admitted between steps while a slot is free and the device
 * memory the server holds plus the observed per-request footprint stays
 * within mem_budget bytes (0 = half of the free device memory at creation).
 * Requests admitted while others are in flight join the same frontier and
 * overlap with them on the device. */
int d4_server_create(const d4_library *lib, uint32_t slots, uint32_t max_height,
                     uint64_t mem_budget, d4_server **out);
void d4_server_destroy(d4_server *s);
/* Queue a request: single != 0 asks for the first solution in the CPU
 * searcher's DFS order (solve(..., true), src/Piece.cpp:66-102), else for
 * all tilings (their count, solve_count, src/Piece.cpp:105-139). */
int d4_server_submit(d4_server *s, const uint8_t *rows, uint32_t height,
                     int single, uint64_t *req_id);
/* Admit waiting requests and advance every in-flight request by one row. */
int d4_server_step(d4_server *s, uint32_t *in_flight, uint32_t *waiting);
/* Take one finished request: returns 1 and fills req_id, count (number of
 * tilings) and, for a single request, its solution in the d4_searcher_
 * solutions format (*len bytes; empty when there is no tiling); returns 0
 * when nothing has finished.  A too-small buffer gives D4_ERANGE and keeps
 * the result queued. */
int d4_server_poll(d4_server *s, uint64_t *req_id, uint64_t *count,
                   uint8_t *sol, size_t cap, size_t *len);
/* Engine statistics of the two frontiers, requests admitted so far, steps
 * at which admission waited for memory, device bytes held. */
int d4_server_stats(const d4_server *s, d4_stats *all, d4_stats *one,
                    uint64_t *admitted, uint64_t *mem_waits,
                    uint64_t *device_bytes);

#ifdef __cplusplus
}
#endif
#endif
