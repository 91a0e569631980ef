"""Shared fixtures: the oracle (CPU restatement of the reference backtracker,
oracle/liboracle.so) and the product's Python mirror (d4-pieces_amd/)."""
import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "d4-pieces_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: runs longer than a few seconds")


class Oracle:
    """ctypes view of oracle/oracle.h (test infrastructure only)."""

    def __init__(self, path):
        L = ctypes.CDLL(path)
        L.or_transform.restype = ctypes.c_uint64
        L.or_transform.argtypes = [ctypes.c_uint64, ctypes.c_uint]
        L.or_canonical.restype = ctypes.c_uint64
        L.or_canonical.argtypes = [ctypes.c_uint64, ctypes.c_uint]
        L.or_enum_free.restype = ctypes.c_size_t
        L.or_enum_free.argtypes = [ctypes.c_uint, ctypes.c_void_p, ctypes.c_size_t]
        L.or_piece_orients.restype = ctypes.c_uint
        L.or_piece_orients.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.or_solve_count.restype = ctypes.c_uint64
        L.or_solve_count.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]
        L.or_solve.restype = ctypes.c_uint64
        L.or_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_size_t,
                               ctypes.POINTER(ctypes.c_size_t)]
        L.or_solve_ex.restype = ctypes.c_uint64
        L.or_solve_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                  ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.or_solve_count_budget.restype = ctypes.c_uint64
        L.or_solve_count_budget.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
        L.or_step_matrix.argtypes = [ctypes.c_uint, ctypes.c_void_p]
        L.or_parse_R.argtypes = [ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint8,
                                 ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p]
        L.or_assemble_R.argtypes = [ctypes.c_uint, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_uint8)]
        self.L = L

    def transform(self, s, t):
        return self.L.or_transform(s, t)

    def canonical(self, s, forms=0xFF):
        return self.L.or_canonical(s, forms)

    def enum_free(self, n):
        buf = (ctypes.c_uint64 * 4096)()
        k = self.L.or_enum_free(n, buf, 4096)
        return list(buf[:k])

    def orients(self, s, sym=0xFF):
        nm = (ctypes.c_uint64 * 8)()
        tr = (ctypes.c_uint * 8)()
        k = self.L.or_piece_orients(s, sym, nm, tr)
        return list(zip(nm[:k], tr[:k]))

    def parse_R(self, H, r, ea):
        rr = (ctypes.c_uint32 * 5)(*r)
        e, c = ctypes.c_uint64(), ctypes.c_uint32()
        nm = (ctypes.c_uint8 * 16)()
        self.L.or_parse_R(H, rr, ea, ctypes.byref(e), ctypes.byref(c), nm)
        return e.value, c.value, list(nm)

    def assemble_R(self, H, empty_area, cnt, nm):
        n = (ctypes.c_uint8 * 16)(*(list(nm) + [0xFF] * (16 - len(nm))))
        r = (ctypes.c_uint32 * 5)()
        ea = ctypes.c_uint8()
        self.L.or_assemble_R(H, empty_area, cnt, n, r, ctypes.byref(ea))
        return tuple(r), ea.value

    @staticmethod
    def _args(shapes, counts, rows):
        sh = (ctypes.c_uint64 * len(shapes))(*shapes)
        cn = (ctypes.c_uint32 * len(shapes))(*counts)
        rw = (ctypes.c_uint8 * max(len(rows), 1))(*rows)
        return sh, cn, rw

    def count(self, shapes, counts, rows, sym=0xFF):
        sh, cn, rw = self._args(shapes, counts, rows)
        return self.L.or_solve_count(sh, cn, len(shapes), sym, rw, len(rows))

    def count_budget(self, shapes, counts, rows, max_nodes, sym=0xFF):
        """(count, complete): complete is False when max_nodes ran out."""
        sh, cn, rw = self._args(shapes, counts, rows)
        done = ctypes.c_int()
        n = self.L.or_solve_count_budget(sh, cn, len(shapes), sym, rw, len(rows), max_nodes,
                                         ctypes.byref(done))
        return n, bool(done.value)

    def solve(self, shapes, counts, rows, single=False, sym=0xFF, enabled=None):
        """enabled: per-piece placement masks (or_solve_ex), None = defaults."""
        sh, cn, rw = self._args(shapes, counts, rows)
        en = (ctypes.c_uint8 * len(shapes))(*enabled) if enabled is not None else None
        need = ctypes.c_size_t(0)
        self.L.or_solve_ex(sh, cn, en, len(shapes), sym, rw, len(rows), int(single), None, 0,
                           ctypes.byref(need))
        buf = (ctypes.c_uint8 * max(need.value, 1))()
        n = self.L.or_solve_ex(sh, cn, en, len(shapes), sym, rw, len(rows), int(single), buf,
                               need.value, ctypes.byref(need))
        data = bytes(buf[:need.value])
        out, i = [], 0
        while i < len(data):
            k = data[i]
            i += 1
            out.append([tuple(data[i + 4 * j:i + 4 * j + 4]) for j in range(k)])
            i += 4 * k
        assert len(out) == n
        return out


@pytest.fixture(scope="session")
def oracle():
    path = os.path.join(ORACLE_DIR, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return Oracle(path)


@pytest.fixture(scope="session")
def d4():
    import d4pieces
    d4pieces.load()
    return d4pieces


@pytest.fixture(scope="session")
def gpu(d4):
    if d4.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return d4
