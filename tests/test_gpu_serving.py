"""Serving (continuous batching): results of requests alone equal their
results in a mix where requests join while others are on the device, and
match the oracle (solve_count / solve(single=true), src/Piece.cpp:66-139)."""
import pytest

import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def p5(gpu):
    return gpu.free_polyominoes(5)


@pytest.fixture(scope="module")
def lib8(gpu, p5):
    return gpu.PieceLibrary(p5, [1] * 12, 0xFF, 8)


def drain(srv, want):
    out = {}
    for _ in range(10000):
        while True:
            r = srv.poll()
            if r is None:
                break
            out[r[0]] = (r[1], r[2])
        if len(out) == want:
            return out
        srv.step()
    raise AssertionError("requests did not finish")


def alone(gpu, lib8, rows, single):
    srv = gpu.Server(lib8, slots=4)
    rid = srv.submit(rows, single)
    return drain(srv, 1)[rid]


def test_mix_equals_alone(gpu, lib8):
    reqs = workloads.request_stream(24, rate=1000.0, seed=4)
    want = [alone(gpu, lib8, rows, single) for _, rows, single in reqs]
    srv = gpu.Server(lib8, slots=64)
    ids = []
    # requests arrive over several steps, so later ones join a running frontier
    for j, (_, rows, single) in enumerate(reqs):
        ids.append(srv.submit(rows, single))
        if j % 3 == 2:
            srv.step()
    got = drain(srv, len(reqs) - 0)
    for rid, w in zip(ids, want):
        assert got[rid] == w


def test_served_results_match_oracle(gpu, lib8, oracle, p5):
    reqs = workloads.request_stream(10, rate=1000.0, seed=7)
    srv = gpu.Server(lib8, slots=16)
    ids = [srv.submit(rows, single) for _, rows, single in reqs]
    got = drain(srv, len(reqs))
    for rid, (_, rows, single) in zip(ids, reqs):
        count, sol = got[rid]
        assert count == oracle.count(p5, [1] * 12, rows)
        if single:
            first = oracle.solve(p5, [1] * 12, rows, single=True)
            assert sol == (first[0] if first else None)


def test_admission_waits_for_memory(gpu, lib8):
    reqs = workloads.request_stream(12, rate=1000.0, seed=11)
    want = [alone(gpu, lib8, rows, single) for _, rows, single in reqs]
    # a budget below two requests' footprint forces admission to wait
    srv = gpu.Server(lib8, slots=64, mem_budget=1 << 20)
    ids = [srv.submit(rows, single) for _, rows, single in reqs]
    got = drain(srv, len(reqs))
    assert [got[i] for i in ids] == want
    assert srv.stats()["mem_waits"] > 0


def test_slot_reuse(gpu, lib8):
    # more requests than slots: slots are freed and reused
    reqs = workloads.request_stream(9, rate=1000.0, seed=12)
    want = [alone(gpu, lib8, rows, single) for _, rows, single in reqs]
    srv = gpu.Server(lib8, slots=2)
    ids = [srv.submit(rows, single) for _, rows, single in reqs]
    got = drain(srv, len(reqs))
    assert [got[i] for i in ids] == want
