"""Python host mirror of the d4 C ABI (include/d4.h).

The product is libd4.so (gfx950 HIP kernels + C++ host engine).  This module
is a thin ctypes layer over it with the reference's names, so that tests and
benchmarks read like the reference's own driver code:

  CudaSearcher     <- struct CudaSearcher       src/searcher_cuda.h:10-38
                      (get_height, size, next_size, search_GPU)
  PieceLibrary     <- g_nme/g_sym + compute_frow_on_cpu + transfer_frow_to_gpu
                      src/frow.h:8-12, src/frow.cpp:39-166
  free_polyominoes <- known_shapes[n]           src/known1.cpp:3-14
  Batch            <- SearcherFactory::run over many board configs,
                      src/searcher.cpp:86-107, as one shared device frontier
  MultiBatch       <- the same with a piece library per instance (the
                      reference's library is a process global, src/frow.h:8-9)
  Solution, expand_solution
                   <- Solution<L> / Step<L> with the cell map
                      (src/Piece.hpp:31-48, src/Piece.cpp:25-54)

There is no fallback: if libd4.so is missing or no HIP device is visible the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libd4.so")

D4_OK, D4_EINVAL, D4_ENOMEM, D4_EDEVICE, D4_ESTATE, D4_ERANGE = 0, -1, -2, -3, -4, -5
MODE_COUNT, MODE_ENUM = 0, 1
FLAG_NO_PRUNE = 1
FLAG_HOST_TABLES = 2
KERNEL_DFS, KERNEL_EXPAND, KERNEL_LEVEL = 0, 1, 2
STEP_GENERAL, STEP_FUSED, STEP_GRAPH = 0, 1, 2
DEDUP_TABLE, DEDUP_PART = 0, 1


class D4Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"d4 error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("rows_done", ctypes.c_uint64),
        ("records_extended", ctypes.c_uint64),
        ("children", ctypes.c_uint64),
        ("unique", ctypes.c_uint64),
        ("pruned", ctypes.c_uint64),
        ("peak_frontier", ctypes.c_uint64),
        ("retries", ctypes.c_uint64),
        ("gpu_ms", ctypes.c_double),
        ("extend_ms", ctypes.c_double),
        ("extend_launches", ctypes.c_uint64),
        ("extend_bytes", ctypes.c_uint64),
        ("compact_ms", ctypes.c_double),
        ("compact_launches", ctypes.c_uint64),
        ("compact_bytes", ctypes.c_uint64),
        ("count_passes", ctypes.c_uint64),
        ("splits", ctypes.c_uint64),
        ("donated", ctypes.c_uint64),
        ("rehashes", ctypes.c_uint64),
        ("sink_growths", ctypes.c_uint64),
        ("host_syncs", ctypes.c_uint64),
        ("dedup_ms", ctypes.c_double),
        ("dedup_launches", ctypes.c_uint64),
        ("dedup_bytes", ctypes.c_uint64),
        ("chunk_reruns", ctypes.c_uint64),
        ("donation_full", ctypes.c_uint64),
        ("fused_steps", ctypes.c_uint64),
        ("fused_fallbacks", ctypes.c_uint64),
        ("graph_replays", ctypes.c_uint64),
        ("host_parked_peak", ctypes.c_uint64),
        ("streamed_steps", ctypes.c_uint64),
        ("stream_passes", ctypes.c_uint64),
        ("part_steps", ctypes.c_uint64),
        ("part_fallbacks", ctypes.c_uint64),
        ("part_remerges", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib: Optional[ctypes.CDLL] = None

# (name, restype, argtypes) of every symbol include/d4.h declares
_P = ctypes.c_void_p
_U8P = ctypes.POINTER(ctypes.c_uint8)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_SZP = ctypes.POINTER(ctypes.c_size_t)
SIGNATURES = [
    ("d4_last_error", ctypes.c_char_p, []),
    ("d4_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("d4_device_synchronize", ctypes.c_int, []),
    ("d4_free_polyominoes", ctypes.c_size_t, [ctypes.c_uint, _U64P, ctypes.c_size_t]),
    ("d4_library_create", ctypes.c_int,
     [_U64P, _U32P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_P)]),
    ("d4_library_create_ex", ctypes.c_int,
     [_U64P, _U32P, ctypes.c_uint32, ctypes.c_uint32, _U8P, ctypes.c_uint32,
      ctypes.POINTER(_P)]),
    ("d4_library_destroy", None, [_P]),
    ("d4_library_info", ctypes.c_int, [_P, _U32P, _U32P]),
    ("d4_library_set_launch", ctypes.c_int,
     [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("d4_library_get_launch", ctypes.c_int, [_P, _U32P, _U32P, _U32P]),
    ("d4_library_set_budget", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_library_set_kernel", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    ("d4_library_set_donation_capacity", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_library_set_step_mode", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_mlib_table_bytes", ctypes.c_uint64, []),
    ("d4_device_atomic_rate", ctypes.c_int, [ctypes.POINTER(ctypes.c_double)]),
    ("d4_mbatch_tables", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]),
    ("d4_library_host_tables", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_uint64]),
    ("d4_library_get_step_mode", ctypes.c_int, [_P, _U32P]),
    ("d4_device_release_cached", ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64)]),
    ("d4_library_set_dedup", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_library_get_dedup", ctypes.c_int, [_P, _U32P]),
    ("d4_searcher_size_bucket", ctypes.c_uint64, [_P, ctypes.c_uint]),
    ("d4_searcher_set_sample", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint64]),
    ("d4_searcher_sample", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                          ctypes.c_void_p]),
    ("d4_searcher_frontier", ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("d4_searcher_next_size_bucket", ctypes.c_uint64, [_P, ctypes.c_uint]),
    ("d4_mbatch_set_donation_capacity", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_mbatch_set_kernel", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_searcher_count_noalloc", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.c_void_p]),
    ("d4_library_get_kernel", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_void_p]),
    ("d4_library_get_budget", ctypes.c_int, [_P, _U32P]),
    ("d4_library_tune", ctypes.c_int,
     [_P, ctypes.c_uint32, _U8P, ctypes.c_uint32, _U32P, _U32P, _U32P, _U32P,
      ctypes.POINTER(ctypes.c_double)]),
    ("d4_searcher_create_ea", ctypes.c_int,
     [_P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_P)]),
    ("d4_searcher_create", ctypes.c_int,
     [_P, _U8P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_P)]),
    ("d4_searcher_destroy", None, [_P]),
    ("d4_searcher_height", ctypes.c_uint32, [_P]),
    ("d4_searcher_size", ctypes.c_uint64, [_P]),
    ("d4_searcher_next_size", ctypes.c_uint64, [_P]),
    ("d4_searcher_step", ctypes.c_int, [_P]),
    ("d4_searcher_run", ctypes.c_int, [_P]),
    ("d4_searcher_set_memory", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_searcher_set_host_memory", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_batch_set_host_memory", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_searcher_count", ctypes.c_int, [_P, _U64P]),
    ("d4_searcher_solutions", ctypes.c_int,
     [_P, ctypes.c_int, _U8P, ctypes.c_size_t, _SZP, _U64P]),
    ("d4_searcher_stats", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("d4_searcher_export", ctypes.c_int, [_P, _P, ctypes.c_size_t, _SZP, _U64P]),
    ("d4_batch_create", ctypes.c_int,
     [_P, ctypes.c_uint32, _U8P, ctypes.c_uint32, _U32P, ctypes.c_uint32, ctypes.c_uint32,
      ctypes.POINTER(_P)]),
    ("d4_batch_destroy", None, [_P]),
    ("d4_batch_run", ctypes.c_int, [_P]),
    ("d4_batch_set_memory", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_batch_counts", ctypes.c_int, [_P, _U64P]),
    ("d4_batch_stats", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("d4_batch_solutions", ctypes.c_int,
     [_P, ctypes.c_uint32, ctypes.c_int, _U8P, ctypes.c_size_t, _SZP, _U64P]),
    ("d4_mbatch_create", ctypes.c_int,
     [ctypes.c_uint32, ctypes.POINTER(_P), _U8P, ctypes.c_uint32, _U32P, ctypes.c_uint32,
      ctypes.POINTER(_P)]),
    ("d4_mbatch_destroy", None, [_P]),
    ("d4_mbatch_set_launch", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    ("d4_mbatch_set_budget_steps", ctypes.c_int, [_P, ctypes.c_uint32]),
    ("d4_mbatch_tune", ctypes.c_int,
     [_P, ctypes.c_uint32, _U32P, _U32P, ctypes.POINTER(ctypes.c_double)]),
    ("d4_mbatch_set_memory", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_mbatch_set_budget", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("d4_mbatch_run", ctypes.c_int, [_P]),
    ("d4_mbatch_instances", ctypes.c_int, [_P, _U64P, _U8P]),
    ("d4_mbatch_counts", ctypes.c_int, [_P, _U64P]),
    ("d4_mbatch_stats", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("d4_mbatch_set_terminals", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint64]),
    ("d4_mbatch_terminals", ctypes.c_int,
     [_P, _U32P, _U64P, _U64P, ctypes.c_size_t, _SZP]),
    ("d4_library_counters", ctypes.c_int, [_P, _U32P, _U32P, _U32P]),
    ("d4_polyominoes", ctypes.c_size_t,
     [ctypes.c_uint, ctypes.c_uint32, _U64P, ctypes.c_size_t]),
    ("d4_naming_create", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
      _U64P, ctypes.POINTER(_P)]),
    ("d4_naming_destroy", None, [_P]),
    ("d4_naming_size", ctypes.c_uint64, [_P]),
    ("d4_naming_size_pieces", ctypes.c_uint64, [_P]),
    ("d4_naming_name_piece", ctypes.c_uint64, [_P, ctypes.c_uint64, ctypes.c_uint64]),
    ("d4_naming_resolve_piece", ctypes.c_int, [_P, ctypes.c_uint64, _U64P, _U64P]),
    ("d4_naming_name", ctypes.c_int, [_P, _U8P, _U64P]),
    ("d4_naming_resolve", ctypes.c_int, [_P, ctypes.c_uint64, _U8P]),
    ("d4_naming_name_counters", ctypes.c_int, [_P, _P, _U64P, ctypes.c_size_t, _U64P]),
    ("d4_ledger_open", ctypes.c_int,
     [ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(_P)]),
    ("d4_ledger_close", None, [_P]),
    ("d4_ledger_should_run", ctypes.c_int,
     [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    ("d4_ledger_log", ctypes.c_int,
     [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    ("d4_ledger_log_many", ctypes.c_int, [_P, _U64P, _U64P, ctypes.c_size_t, _U8P]),
    ("d4_ledger_after_run", ctypes.c_int,
     [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
    ("d4_ledger_progress", ctypes.c_int, [_P, _U64P, _U64P]),
    ("d4_ledger_names", ctypes.c_int, [_P, _U64P, ctypes.c_size_t]),
    ("d4_server_create", ctypes.c_int,
     [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_P)]),
    ("d4_server_destroy", None, [_P]),
    ("d4_server_submit", ctypes.c_int,
     [_P, _U8P, ctypes.c_uint32, ctypes.c_int, _U64P]),
    ("d4_server_step", ctypes.c_int, [_P, _U32P, _U32P]),
    ("d4_server_poll", ctypes.c_int, [_P, _U64P, _U64P, _U8P, ctypes.c_size_t, _SZP]),
    ("d4_server_stats", ctypes.c_int,
     [_P, ctypes.POINTER(Stats), ctypes.POINTER(Stats), _U64P, _U64P, _U64P]),
]


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libd4.so (built by `make -C d4-pieces_amd`); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `make -C d4-pieces_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != D4_OK:
        msg = load().d4_last_error()
        raise D4Error(rc, msg.decode() if msg else "")


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load().d4_device_count(ctypes.byref(n)))
    return n.value


def atomic_rate() -> float:
    """Measured random 64-bit device-scope atomic adds per second."""
    v = ctypes.c_double()
    _check(load().d4_device_atomic_rate(ctypes.byref(v)))
    return v.value


def release_cached() -> int:
    """Return the allocator's cached device memory to the runtime; bytes."""
    v = ctypes.c_uint64()
    _check(load().d4_device_release_cached(ctypes.byref(v)))
    return v.value


def synchronize() -> None:
    _check(load().d4_device_synchronize())


def free_polyominoes(n: int) -> List[int]:
    """Canonical free polyominoes with n cells, ascending (known_shapes[n])."""
    lib = load()
    total = lib.d4_free_polyominoes(n, None, 0)
    buf = (ctypes.c_uint64 * max(total, 1))()
    lib.d4_free_polyominoes(n, buf, total)
    return list(buf[:total])


def _u64arr(v):
    return (ctypes.c_uint64 * len(v))(*v)


def _u32arr(v):
    return (ctypes.c_uint32 * len(v))(*v)


def _u8arr(v):
    return (ctypes.c_uint8 * max(len(v), 1))(*v)


Step = Tuple[int, int, int, int]  # (piece_id, trs_id, y, x)


def normalize8(v: int) -> int:
    """Translate a shape (8-wide encoding, bit = row*8 + col) to the top-left
    corner (Shape::normalize)."""
    if not v:
        return 0
    while not v & 0xFF:
        v >>= 8
    while not v & 0x0101010101010101:
        v >>= 1
    return v


def transform8(v: int, t: int) -> int:
    """D4 transform t in the reference's order (src/Shape.cpp:6-40: bit 2 =
    transpose, bit 0 = flip X, bit 1 = flip Y, applied to the input's
    coordinates), normalised."""
    sw, fx, fy = t & 4, t & 1, t & 2
    out = 0
    for b in range(64):
        if v >> b & 1:
            y, x = divmod(b, 8)
            yy, xx = (7 - y if fy else y), (7 - x if fx else x)
            r, c = (xx, yy) if sw else (yy, xx)
            out |= 1 << (r * 8 + c)
    return normalize8(out)


# make_step's matrix per transform (src/Piece.cpp:41-54)
_STEP_MATRIX = [(1, 0, 0, 1), (-1, 0, 0, 1), (1, 0, 0, -1), (-1, 0, 0, -1),
                (0, 1, 1, 0), (0, -1, 1, 0), (0, 1, -1, 0), (0, -1, -1, 0)]


class FullStep:
    """Step<L> (src/Piece.hpp:31-36): piece id, transform id, the transform's
    2x2 matrix a, b, c, d, the translation x, y and the placed shape."""
    __slots__ = ("piece_id", "trs_id", "a", "b", "c", "d", "x", "y", "shape")

    def __init__(self, piece_id, trs_id, a, b, c, d, x, y, shape):
        self.piece_id, self.trs_id = piece_id, trs_id
        self.a, self.b, self.c, self.d, self.x, self.y = a, b, c, d, x, y
        self.shape = shape

    def __repr__(self):
        return (f"Step(piece={self.piece_id}, trs={self.trs_id}, x={self.x}, y={self.y}, "
                f"shape={self.shape:#x})")


class Solution:
    """Solution<L> (src/Piece.hpp:38-48, src/Piece.cpp:25-39): the steps and
    map[row][col] = index of the step covering the cell, -1 for none
    (8 columns, one row per board row)."""

    def __init__(self, steps: List[FullStep], height: int):
        self.steps = steps
        self.map = [[next((i for i, st in enumerate(steps) if st.shape >> (r * 8 + c) & 1), -1)
                     for c in range(8)] for r in range(height)]


def expand_solution(shapes: Sequence[int], raw: Sequence[Step], height: int,
                    sym: int = 0xFF) -> Solution:
    """A solution as returned by solutions() ((piece_id, trs_id, y, x) per
    step) with the steps' matrices, placed shapes and the cell map.  trs_id
    counts the transforms enabled in `sym` (Piece.cpp:12-22); the placed
    shape is that transform's, while the matrix is make_step's for trs_id
    itself (Piece.cpp:41-54 switches on the placement index, so under a
    partial sym mask it is not the placed transform's matrix)."""
    enabled = [t for t in range(8) if sym >> t & 1]
    steps = []
    for pid, trs, y, x in raw:
        shape = transform8(shapes[pid], enabled[trs]) << (8 * y + x)
        steps.append(FullStep(pid, trs, *_STEP_MATRIX[trs], x, y, shape))
    return Solution(steps, height)


def decode_solutions(buf: bytes) -> List[List[Step]]:
    out, i = [], 0
    while i < len(buf):
        n = buf[i]
        i += 1
        out.append([tuple(buf[i + 4 * j: i + 4 * j + 4]) for j in range(n)])
        i += 4 * n
    return out


class PieceLibrary:
    """Piece set + placement tables on the current device."""

    def __init__(self, shapes: Sequence[int], counts: Optional[Sequence[int]] = None,
                 sym: int = 0xFF, width: int = 8, enabled: Optional[Sequence[int]] = None):
        """enabled[i]: piece i's placement mask (bit p = the p-th transform sym
        admits); None = every distinct placement."""
        lib = load()
        counts = list(counts) if counts is not None else [1] * len(shapes)
        self.shapes, self.counts, self.sym, self.width = list(shapes), counts, sym, width
        h = _P()
        if enabled is None:
            _check(lib.d4_library_create(_u64arr(self.shapes), _u32arr(counts), len(shapes),
                                         sym, width, ctypes.byref(h)))
        else:
            _check(lib.d4_library_create_ex(_u64arr(self.shapes), _u32arr(counts), len(shapes),
                                            sym, _u8arr(list(enabled)), width, ctypes.byref(h)))
        self._h = h
        nc, mh = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.d4_library_info(h, ctypes.byref(nc), ctypes.byref(mh)))
        self.n_cand, self.max_height = nc.value, mh.value

    def set_launch(self, fit_rows: int, threads: int, refill: int = 16) -> None:
        _check(load().d4_library_set_launch(self._h, fit_rows, threads, refill))

    def set_budget(self, steps: int) -> None:
        """DFS steps per work item before a lane donates the rest (0: never)."""
        _check(load().d4_library_set_budget(self._h, steps))

    def get_budget(self) -> int:
        v = ctypes.c_uint32()
        _check(load().d4_library_get_budget(self._h, ctypes.byref(v)))
        return v.value

    def set_kernel(self, kernel: int, tail_rounds: int = 256) -> None:
        """Count-mode row-step kernel: KERNEL_LEVEL (default), KERNEL_EXPAND or
        KERNEL_DFS; tail_rounds: k_expand rounds after the records run out
        before a wave's partial rows move to a follow-up launch (0: never)."""
        _check(load().d4_library_set_kernel(self._h, kernel, tail_rounds))

    def set_donation_capacity(self, items: int) -> None:
        """Items per donation buffer (default 1 Mi)."""
        _check(load().d4_library_set_donation_capacity(self._h, items))

    def host_tables(self) -> bytes:
        """This library's multi-library batch tables built on the host."""
        n = load().d4_mlib_table_bytes()
        buf = ctypes.create_string_buffer(n)
        _check(load().d4_library_host_tables(self._h, buf, n))
        return buf.raw

    def set_step_mode(self, mode: int) -> None:
        """How count-mode level row steps are issued: STEP_GENERAL (host round
        trips between the parts), STEP_FUSED (default: one queued chain, one
        round trip) or STEP_GRAPH (that chain replayed from a HIP graph)."""
        _check(load().d4_library_set_step_mode(self._h, mode))

    def set_dedup(self, mode: int) -> None:
        """Dedup of count-mode level row steps: DEDUP_PART (default: hash
        buckets merged in LDS) or DEDUP_TABLE (one table in device memory)."""
        _check(load().d4_library_set_dedup(self._h, mode))

    def get_dedup(self) -> int:
        m = ctypes.c_uint32()
        _check(load().d4_library_get_dedup(self._h, ctypes.byref(m)))
        return m.value

    def get_step_mode(self) -> int:
        m = ctypes.c_uint32()
        _check(load().d4_library_get_step_mode(self._h, ctypes.byref(m)))
        return m.value

    def get_kernel(self) -> Tuple[int, int]:
        k, t = ctypes.c_uint32(), ctypes.c_uint32()
        _check(load().d4_library_get_kernel(self._h, ctypes.byref(k), ctypes.byref(t)))
        return k.value, t.value

    def get_launch(self) -> Tuple[int, int, int]:
        f, t, r = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(load().d4_library_get_launch(self._h, ctypes.byref(f), ctypes.byref(t),
                                            ctypes.byref(r)))
        return f.value, t.value, r.value

    def tune(self, boards: Sequence[Sequence[int]]) -> dict:
        """Measured autotune of the extend launch shape on these boards."""
        flat, stride, heights = _flatten(boards)
        f, t, r, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        _check(load().d4_library_tune(self._h, len(boards), _u8arr(flat), stride,
                                      _u32arr(heights), ctypes.byref(f), ctypes.byref(t),
                                      ctypes.byref(r), ctypes.byref(ms)))
        return {"fit_rows": f.value, "threads": t.value, "refill": r.value,
                "budget": self.get_budget(), "extend_ms": ms.value}

    def counter_layout(self) -> List[Tuple[int, int]]:
        """(bit offset, width) of each piece's use counter in packed counters."""
        off, wid, n = (ctypes.c_uint32 * 64)(), (ctypes.c_uint32 * 64)(), ctypes.c_uint32()
        _check(load().d4_library_counters(self._h, off, wid, ctypes.byref(n)))
        return [(off[i], wid[i]) for i in range(n.value)]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_library_destroy(self._h)
            self._h = None


def _flatten(boards):
    stride = max([len(b) for b in boards] + [1])
    flat = []
    for b in boards:
        flat.extend(list(b) + [0] * (stride - len(b)))
    return flat, stride, [len(b) for b in boards]


def _rows_arg(rows: Sequence[int]):
    return _u8arr(list(rows))


def count_noalloc(lib: "PieceLibrary", rows: Sequence[int], prune: bool = True) -> int:
    """Tilings of one board counted in the device's fixed scratch, with no
    device allocation (d4_searcher_count_noalloc)."""
    rows = list(rows)
    out = ctypes.c_uint64()
    _check(load().d4_searcher_count_noalloc(lib._h, _rows_arg(rows), len(rows),
                                            0 if prune else FLAG_NO_PRUNE, ctypes.byref(out)))
    return out.value


class CudaSearcher:
    """One board, advanced one row per search_GPU() (CudaSearcher)."""

    def __init__(self, library: PieceLibrary, rows: Optional[Sequence[int]] = None,
                 empty_area: Optional[int] = None, mode: int = MODE_COUNT,
                 prune: bool = True):
        lib = load()
        self.library = library
        flags = 0 if prune else FLAG_NO_PRUNE
        h = _P()
        if empty_area is not None:
            _check(lib.d4_searcher_create_ea(library._h, empty_area, mode, flags,
                                             ctypes.byref(h)))
        else:
            rows = list(rows or [])
            _check(lib.d4_searcher_create(library._h, _rows_arg(rows), len(rows), mode,
                                          flags, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_searcher_destroy(self._h)
            self._h = None

    def get_height(self) -> int:
        return load().d4_searcher_height(self._h)

    def set_sample(self, height: int, cap: int) -> None:
        """Keep up to cap frontier states at `height` rows left as the search
        passes that height (sample() returns them)."""
        _check(load().d4_searcher_set_sample(self._h, height, cap))

    def sample(self) -> List[Tuple[int, int]]:
        n = ctypes.c_uint64()
        _check(load().d4_searcher_sample(self._h, None, None, 0, ctypes.byref(n)))
        keys = (ctypes.c_uint64 * max(n.value, 1))()
        cnts = (ctypes.c_uint64 * max(n.value, 1))()
        _check(load().d4_searcher_sample(self._h, keys, cnts, n.value, ctypes.byref(n)))
        return list(zip(keys[:n.value], cnts[:n.value]))

    def frontier(self, first: int = 0, cap: int = 1 << 16) -> List[Tuple[int, int]]:
        """(key, tiling count) of frontier states first .. first + cap - 1."""
        n = ctypes.c_uint64()
        _check(load().d4_searcher_frontier(self._h, first, 0, None, None, ctypes.byref(n)))
        k = max(0, min(cap, n.value - first))
        keys = (ctypes.c_uint64 * max(k, 1))()
        cnts = (ctypes.c_uint64 * max(k, 1))()
        _check(load().d4_searcher_frontier(self._h, first, k, keys, cnts, ctypes.byref(n)))
        return list(zip(keys[:k], cnts[:k]))

    def size(self, i: Optional[int] = None) -> int:
        """Records in the frontier (size()), or in bucket i, the records whose
        current row has open cells i (size(unsigned i))."""
        if i is None:
            return load().d4_searcher_size(self._h)
        return load().d4_searcher_size_bucket(self._h, i)

    def next_size(self, i: Optional[int] = None) -> int:
        """Upper bound on the next step's children, in total or for bucket i."""
        if i is None:
            return load().d4_searcher_next_size(self._h)
        return load().d4_searcher_next_size_bucket(self._h, i)

    def set_memory(self, nbytes: int) -> None:
        """Count mode: device bytes run() may use (frontier parts beyond it
        are parked and finished one after the other)."""
        _check(load().d4_searcher_set_memory(self._h, nbytes))

    def set_host_memory(self, nbytes: int) -> None:
        """Count mode: pinned host bytes parked frontier parts may take (0:
        they stay in device memory)."""
        _check(load().d4_searcher_set_host_memory(self._h, nbytes))

    def search_GPU(self) -> None:
        _check(load().d4_searcher_step(self._h))

    def run(self) -> "CudaSearcher":
        _check(load().d4_searcher_run(self._h))
        return self

    def count(self) -> int:
        v = ctypes.c_uint64()
        _check(load().d4_searcher_count(self._h, ctypes.byref(v)))
        return v.value

    def solutions(self, single: bool = False) -> List[List[Step]]:
        lib = load()
        ln, ns = ctypes.c_size_t(), ctypes.c_uint64()
        _check(lib.d4_searcher_solutions(self._h, int(single), None, 0, ctypes.byref(ln),
                                         ctypes.byref(ns)))
        buf = (ctypes.c_uint8 * max(ln.value, 1))()
        _check(lib.d4_searcher_solutions(self._h, int(single), buf, ln.value,
                                         ctypes.byref(ln), ctypes.byref(ns)))
        return decode_solutions(bytes(buf[:ln.value]))

    def stats(self) -> dict:
        st = Stats()
        _check(load().d4_searcher_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def export(self):
        """The frontier as the reference's RX records: list of (ea, R words
        (5 ints), count), ordered by (ea, R)."""
        lib = load()
        n = ctypes.c_size_t()
        _check(lib.d4_searcher_export(self._h, None, 0, ctypes.byref(n), None))
        buf = (ctypes.c_uint32 * (6 * max(n.value, 1)))()
        cnt = (ctypes.c_uint64 * max(n.value, 1))()
        _check(lib.d4_searcher_export(self._h, buf, n.value, ctypes.byref(n), cnt))
        return [(buf[6 * i + 5] & 0xFF, tuple(buf[6 * i:6 * i + 5]), cnt[i])
                for i in range(n.value)]


class Batch:
    """Many boards sharing one library, advanced together in one frontier."""

    def __init__(self, library: PieceLibrary, boards: Sequence[Sequence[int]],
                 mode: int = MODE_COUNT, prune: bool = True):
        lib = load()
        self.library = library
        self.n = len(boards)
        flat, stride, heights = _flatten(boards)
        h = _P()
        _check(lib.d4_batch_create(library._h, self.n, _u8arr(flat), stride, _u32arr(heights),
                                   mode, 0 if prune else FLAG_NO_PRUNE, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_batch_destroy(self._h)
            self._h = None

    def set_memory(self, nbytes: int) -> None:
        _check(load().d4_batch_set_memory(self._h, nbytes))

    def set_host_memory(self, nbytes: int) -> None:
        _check(load().d4_batch_set_host_memory(self._h, nbytes))

    def run(self) -> "Batch":
        _check(load().d4_batch_run(self._h))
        return self

    def counts(self) -> List[int]:
        buf = (ctypes.c_uint64 * self.n)()
        _check(load().d4_batch_counts(self._h, buf))
        return list(buf)

    def solutions(self, i: int, single: bool = False) -> List[List[Step]]:
        lib = load()
        ln, ns = ctypes.c_size_t(), ctypes.c_uint64()
        _check(lib.d4_batch_solutions(self._h, i, int(single), None, 0, ctypes.byref(ln),
                                      ctypes.byref(ns)))
        buf = (ctypes.c_uint8 * max(ln.value, 1))()
        _check(lib.d4_batch_solutions(self._h, i, int(single), buf, ln.value,
                                      ctypes.byref(ln), ctypes.byref(ns)))
        return decode_solutions(bytes(buf[:ln.value]))

    def stats(self) -> dict:
        st = Stats()
        _check(load().d4_batch_stats(self._h, ctypes.byref(st)))
        return st.as_dict()


class MultiBatch:
    """Instances with their own piece libraries and board widths, advanced
    together in one frontier (d4_mbatch_*); count mode."""

    def __init__(self, libraries: Sequence[PieceLibrary], boards: Sequence[Sequence[int]],
                 prune: bool = True, host_tables: bool = False, kernel: Optional[int] = None):
        lib = load()
        if len(libraries) != len(boards):
            raise ValueError("one library per board")
        self.libraries = list(libraries)  # keep the handles alive
        self.n = len(boards)
        flat, stride, heights = _flatten(boards)
        handles = (_P * self.n)(*[L._h for L in libraries])
        h = _P()
        flags = (0 if prune else FLAG_NO_PRUNE) | (FLAG_HOST_TABLES if host_tables else 0)
        _check(lib.d4_mbatch_create(self.n, handles, _u8arr(flat), stride, _u32arr(heights),
                                    flags, ctypes.byref(h)))
        self._h = h
        if kernel is not None:
            self.set_kernel(kernel)

    def tables(self, inst: int) -> bytes:
        """Instance inst's library tables as the batch holds them on the device."""
        n = load().d4_mlib_table_bytes()
        buf = ctypes.create_string_buffer(n)
        _check(load().d4_mbatch_tables(self._h, inst, buf, n))
        return buf.raw

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_mbatch_destroy(self._h)
            self._h = None

    def set_launch(self, threads: int, refill: int = 16) -> None:
        _check(load().d4_mbatch_set_launch(self._h, threads, refill))

    def set_kernel(self, kernel: int) -> None:
        """Row-step kernel: KERNEL_LEVEL (default) or KERNEL_DFS (the one
        set_launch, tune, set_budget_steps and set_donation_capacity shape)."""
        _check(load().d4_mbatch_set_kernel(self._h, kernel))

    def set_memory(self, nbytes: int) -> None:
        _check(load().d4_mbatch_set_memory(self._h, nbytes))

    def tune(self, sample: int = 256) -> dict:
        """Measured choice of workgroup size and refill threshold on the first
        `sample` instances (applied to this batch)."""
        t, r, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        _check(load().d4_mbatch_tune(self._h, sample, ctypes.byref(t), ctypes.byref(r),
                                     ctypes.byref(ms)))
        return {"threads": t.value, "refill": r.value, "extend_ms": ms.value}

    def set_budget_steps(self, steps: int) -> None:
        """DFS steps per work item of a row's main launch before a lane donates
        the rest (default 4096; 0: never)."""
        _check(load().d4_mbatch_set_budget_steps(self._h, steps))

    def set_donation_capacity(self, items: int) -> None:
        """Items per donation buffer (default 1 Mi)."""
        _check(load().d4_mbatch_set_donation_capacity(self._h, items))

    def set_budget(self, max_states: int) -> None:
        """States one instance may hold in a frontier before it is dropped
        (reported unfinished); 0 = no limit."""
        _check(load().d4_mbatch_set_budget(self._h, max_states))

    def instances(self) -> Tuple[List[int], List[bool]]:
        """(peak states per instance, finished per instance)."""
        peak = (ctypes.c_uint64 * self.n)()
        fin = (ctypes.c_uint8 * self.n)()
        _check(load().d4_mbatch_instances(self._h, peak, fin))
        return list(peak), [bool(x) for x in fin]

    def run(self) -> "MultiBatch":
        _check(load().d4_mbatch_run(self._h))
        return self

    def counts(self) -> List[int]:
        buf = (ctypes.c_uint64 * self.n)()
        _check(load().d4_mbatch_counts(self._h, buf))
        return list(buf)

    def stats(self) -> dict:
        st = Stats()
        _check(load().d4_mbatch_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def set_terminals(self, keep: bool = True, initial_entries: int = 0) -> None:
        """Keep the terminal states instances reach (for terminals()); the
        device buffer starts with room for initial_entries (0: default)."""
        _check(load().d4_mbatch_set_terminals(self._h, 1 if keep else 0, initial_entries))

    def terminals(self) -> Tuple[List[int], List[int], List[int]]:
        """(instance, packed piece counters, tilings) per distinct terminal
        state of a finished instance, sorted."""
        lib = load()
        n = ctypes.c_size_t()
        _check(lib.d4_mbatch_terminals(self._h, None, None, None, 0, ctypes.byref(n)))
        k = max(n.value, 1)
        inst, ctr, til = (ctypes.c_uint32 * k)(), (ctypes.c_uint64 * k)(), (ctypes.c_uint64 * k)()
        _check(lib.d4_mbatch_terminals(self._h, inst, ctr, til, n.value, ctypes.byref(n)))
        m = n.value
        return list(inst[:m]), list(ctr[:m]), list(til[:m])


def polyominoes(n: int, sym: int = 0xFF) -> List[int]:
    """Canonical polyominoes with n cells under the transforms in sym,
    ascending: sym 0xFF = known_shapes[n] (src/known1.cpp), sym 0x69 =
    known_C_shapes[n] (src/known2.cpp)."""
    lib = load()
    total = lib.d4_polyominoes(n, sym, None, 0)
    buf = (ctypes.c_uint64 * max(total, 1))()
    lib.d4_polyominoes(n, sym, buf, total)
    return list(buf[:total])


NO_NAME = (1 << 64) - 1


class Naming:
    """Naming (src/naming.hpp:15-71): the piece sets of between min_n and
    max_n pieces of sizes min_m..max_m, at most one of each, that cover tgt
    cells; arr[m] lists the pieces of size m.  Names are dense ranks in
    [0, size())."""

    def __init__(self, min_m: int, max_m: int, min_n: int, max_n: int, tgt: int,
                 arr: Sequence[Sequence[int]]):
        self.min_m, self.max_m, self.min_n, self.max_n, self.tgt = min_m, max_m, min_n, max_n, tgt
        self.arr = [list(arr[m]) if m < len(arr) and arr[m] else [] for m in range(max_m + 1)]
        sizes = [len(self.arr[m]) for m in range(max_m + 1)]
        h = _P()
        _check(load().d4_naming_create(min_m, max_m, min_n, max_n, tgt, _u64arr(sizes),
                                       ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_naming_destroy(self._h)
            self._h = None

    def size(self) -> int:
        return load().d4_naming_size(self._h)

    def size_pieces(self) -> int:
        return load().d4_naming_size_pieces(self._h)

    def name_piece(self, m: int, i: int) -> int:
        return load().d4_naming_name_piece(self._h, m, i)

    def resolve_piece(self, nm: int) -> Tuple[int, int]:
        """(m, i) of piece id nm; (0, 0) when out of range, as the reference."""
        m, i = ctypes.c_uint64(), ctypes.c_uint64()
        rc = load().d4_naming_resolve_piece(self._h, nm, ctypes.byref(m), ctypes.byref(i))
        if rc == D4_ERANGE:
            return 0, 0
        _check(rc)
        return m.value, i.value

    def pieces(self) -> List[int]:
        """Every piece's shape, in piece-id order."""
        return [v for m in range(self.min_m, self.max_m + 1) for v in self.arr[m]]

    def name_used(self, used: Sequence[int]) -> Optional[int]:
        out = ctypes.c_uint64()
        _check(load().d4_naming_name(self._h, _u8arr(list(used)), ctypes.byref(out)))
        return None if out.value == NO_NAME else out.value

    def name(self, func) -> Optional[int]:
        """Naming::name: func(m, i) is true iff the i-th piece of size m is
        in the set; None when the set is not named."""
        used = [1 if func(m, i) else 0
                for m in range(self.min_m, self.max_m + 1) for i in range(len(self.arr[m]))]
        return self.name_used(used)

    def resolve_used(self, nm: int) -> Optional[List[int]]:
        used = (ctypes.c_uint8 * max(self.size_pieces(), 1))()
        rc = load().d4_naming_resolve(self._h, nm, used)
        if rc == D4_ERANGE:
            return None
        _check(rc)
        return list(used[:self.size_pieces()])

    def resolve(self, nm: int, func) -> bool:
        """Naming::resolve: calls func(m, i) for each piece of set nm;
        False when nm >= size()."""
        used = self.resolve_used(nm)
        if used is None:
            return False
        for pid, u in enumerate(used):
            if u:
                func(*self.resolve_piece(pid))
        return True

    def name_counters(self, library: "PieceLibrary", counters: Sequence[int]) -> List[int]:
        """Names (NO_NAME where none) of terminal states' packed counters of
        a search over `library`, whose piece p is piece id p."""
        n = len(counters)
        out = (ctypes.c_uint64 * max(n, 1))()
        _check(load().d4_naming_name_counters(self._h, library._h, _u64arr(list(counters)) if n
                                              else None, n, out))
        return list(out[:n])


LEDGER_VERIFY, LEDGER_COUNT, LEDGER_BIT, LEDGER_BYTE = 0, 1, 2, 3


class Ledger:
    """Where names are logged per board config (d4_ledger_*): Verify,
    AtomicCount, or the FileBit / FileByte mmap layout of src/search.cpp."""

    def __init__(self, kind: int, n_configs: int, n_names: int, path: Optional[str] = None):
        h = _P()
        _check(load().d4_ledger_open(kind, path.encode() if path else None, n_configs, n_names,
                                     ctypes.byref(h)))
        self._h, self.kind, self.n_configs, self.n_names = h, kind, n_configs, n_names

    def close(self) -> None:
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_ledger_close(self._h)
            self._h = None

    __del__ = close

    def should_run(self, config: int, shape: int) -> bool:
        r = ctypes.c_int()
        _check(load().d4_ledger_should_run(self._h, config, shape, ctypes.byref(r)))
        return bool(r.value)

    def log(self, config: int, name: int) -> bool:
        r = ctypes.c_int()
        _check(load().d4_ledger_log(self._h, config, name, ctypes.byref(r)))
        return bool(r.value)

    def log_many(self, configs: Sequence[int], names: Sequence[int]) -> List[bool]:
        n = len(configs)
        fresh = (ctypes.c_uint8 * max(n, 1))()
        if n:
            _check(load().d4_ledger_log_many(self._h, _u64arr(list(configs)),
                                             _u64arr(list(names)), n, fresh))
        return [bool(x) for x in fresh[:n]]

    def after_run(self, config: int, shape: int, count: int, ms: int) -> None:
        _check(load().d4_ledger_after_run(self._h, config, shape, count, ms))

    def progress(self) -> Tuple[int, int]:
        w, c = ctypes.c_uint64(), ctypes.c_uint64()
        _check(load().d4_ledger_progress(self._h, ctypes.byref(w), ctypes.byref(c)))
        return w.value, c.value

    def names(self) -> List[int]:
        buf = (ctypes.c_uint64 * max(self.n_names, 1))()
        _check(load().d4_ledger_names(self._h, buf, self.n_names))
        return list(buf[:self.n_names])


class Server:
    """Continuous-batching request loop (d4_server_*)."""

    def __init__(self, library: PieceLibrary, slots: int = 256, max_height: int = 8,
                 mem_budget: int = 0):
        self.library = library
        h = _P()
        _check(load().d4_server_create(library._h, slots, max_height, mem_budget,
                                       ctypes.byref(h)))
        self._h = h
        self._buf = (ctypes.c_uint8 * 4096)()

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.d4_server_destroy(self._h)
            self._h = None

    def submit(self, rows: Sequence[int], single: bool) -> int:
        rid = ctypes.c_uint64()
        rows = list(rows)
        _check(load().d4_server_submit(self._h, _u8arr(rows), len(rows), int(single),
                                       ctypes.byref(rid)))
        return rid.value

    def step(self) -> Tuple[int, int]:
        fl, wt = ctypes.c_uint32(), ctypes.c_uint32()
        _check(load().d4_server_step(self._h, ctypes.byref(fl), ctypes.byref(wt)))
        return fl.value, wt.value

    def poll(self):
        """One finished request as (req_id, count, solution or None), or None."""
        rid, cnt, ln = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_size_t()
        rc = load().d4_server_poll(self._h, ctypes.byref(rid), ctypes.byref(cnt), self._buf,
                                   len(self._buf), ctypes.byref(ln))
        if rc == D4_ERANGE:
            self._buf = (ctypes.c_uint8 * (2 * len(self._buf)))()
            return self.poll()
        if rc < 0:
            _check(rc)
        if rc == 0:
            return None
        sols = decode_solutions(bytes(self._buf[:ln.value]))
        return rid.value, cnt.value, (sols[0] if sols else None)

    def stats(self) -> dict:
        a, o = Stats(), Stats()
        adm, mw, db = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(load().d4_server_stats(self._h, ctypes.byref(a), ctypes.byref(o),
                                      ctypes.byref(adm), ctypes.byref(mw), ctypes.byref(db)))
        return {"all": a.as_dict(), "one": o.as_dict(), "admitted": adm.value,
                "mem_waits": mw.value, "device_bytes": db.value}
