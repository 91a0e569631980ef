"""The search driver over board configurations (src/search.cpp:25-320,
src/searcher.cpp:15-107, src/searcher.hpp:16-75): every configuration of a
board file is tiled with the pieces of a Naming, each used at most once;
every distinct piece set that tiles a configuration is named and logged in a
ledger (Verify, AtomicCount, FileBit, FileByte).

The reference searches one configuration per CudaSearcher and leaves the
consumer of its terminal records commented out (src/searcher.cpp:38-65).
Here all configurations that still need a run go through one multi-instance
frontier on the device (d4_mbatch_*, one instance per configuration);
terminal states arrive deduplicated per (configuration, piece counters),
are named in bulk (d4_naming_name_counters) and logged (d4_ledger_log_many).
"""
import time
from typing import List, Optional, Sequence

import d4pieces as d4
import workloads

SYM_FREE, SYM_C = 0xFF, 0x69  # the reference's g_sym: free, or C (env C set)


def known_shapes(sym: int = SYM_FREE) -> List[List[int]]:
    """known_shapes / known_C_shapes (src/known.hpp:9-12) indexed by size."""
    return [[]] + [d4.polyominoes(m, sym) for m in range(1, 9)]


def shape_of(rows: Sequence[int], width: int = 8) -> int:
    v = 0
    for r, bits in enumerate(rows):
        v |= bits << (width * r)
    return v


class SearcherFactory:
    """SearcherFactory (src/searcher.hpp:45-75) with its ledger."""
    kind = d4.LEDGER_VERIFY

    def __init__(self, naming: d4.Naming, n_configs: int, path: Optional[str] = None,
                 sym: int = SYM_FREE, width: int = 8):
        self.naming = naming
        self.n_configs = n_configs
        self.ledger = d4.Ledger(self.kind, n_configs, naming.size(), path)
        pieces = naming.pieces()
        self.library = d4.PieceLibrary(pieces, [1] * len(pieces), sym, width)
        self.width = width
        self.zero = []           # configs no named set tiles (the reference prints an error)
        self.configs_issued = 0

    def should_run(self, i: int, shape: int) -> bool:
        return self.ledger.should_run(i, shape)

    def after_run(self, i: int, shape: int, cnt: int, ms: int) -> None:
        if not cnt:
            self.zero.append(i)
        self.ledger.after_run(i, shape, cnt, ms)

    def work_done(self) -> int:
        return self.ledger.progress()[0]

    def configs_done(self) -> int:
        return self.ledger.progress()[1]

    def run(self, configs: Sequence[Sequence[int]], batch: int = 8192) -> dict:
        """SearcherFactory::run (src/searcher.cpp:86-107) over row-mask
        configs (index = config index); returns per-run figures."""
        shapes = [shape_of(c, self.width) for c in configs]
        todo = [i for i in range(len(configs)) if self.should_run(i, shapes[i])]
        self.configs_issued += len(todo)
        out = {"configs": len(todo), "terminals": 0, "logged": 0, "gpu_s": 0.0}
        for k in range(0, len(todo), batch):
            part = todo[k:k + batch]
            t0 = time.perf_counter()
            b = d4.MultiBatch([self.library] * len(part), [configs[i] for i in part])
            b.set_terminals(True)
            b.run()
            inst, ctr, _ = b.terminals()
            names = self.naming.name_counters(self.library, ctr)
            fresh = self.ledger.log_many([part[j] for j in inst], names)
            dt = time.perf_counter() - t0
            cnt = [0] * len(part)
            for j, f in zip(inst, fresh):
                cnt[j] += f
            ms = max(1, int(round(1000 * dt / len(part))))
            for j, i in enumerate(part):
                self.after_run(i, shapes[i], cnt[j], ms)
            out["terminals"] += len(inst)
            out["logged"] += sum(fresh)
            out["gpu_s"] += dt
        return out


class Verify(SearcherFactory):
    """S=v: counts every named terminal set (src/search.cpp:25-35)."""
    kind = d4.LEDGER_VERIFY


class AtomicCount(SearcherFactory):
    """S=ac: per name, the number of configs it tiles (src/search.cpp:37-54)."""
    kind = d4.LEDGER_COUNT


class FileBit(SearcherFactory):
    """S=fb: one bit per (name, config) in an mmap'd file (search.cpp:200-226)."""
    kind = d4.LEDGER_BIT


class FileByte(SearcherFactory):
    """S=fB: one byte per (name, config) in an mmap'd file (search.cpp:176-198)."""
    kind = d4.LEDGER_BYTE


FACTORIES = {"v": Verify, "ac": AtomicCount, "fb": FileBit, "fB": FileByte}


def main(board_text: str, min_m: int, max_m: int, min_n: int, max_n: int, S: str,
         path: Optional[str] = None, sym_C: bool = False, batch: int = 8192):
    """src/search.cpp:248-320: the board file's configs, a Naming over the
    known shapes of sizes min_m..max_m covering the configs' free cells, and
    the factory named by S.  Returns (factory, run figures)."""
    count, configs = workloads.board_configs(board_text)
    if count != len(configs):
        raise ValueError("board file: configuration count mismatch")
    tgt = bin(shape_of(configs[0])).count("1") if configs else 0
    sym = SYM_C if sym_C else SYM_FREE
    nm = d4.Naming(min_m, max_m, min_n, max_n, tgt, known_shapes(sym))
    if S not in FACTORIES:
        raise ValueError("S must be one of v|ac|fb|fB")
    sf = FACTORIES[S](nm, len(configs), path, sym)
    return sf, sf.run(configs, batch)
