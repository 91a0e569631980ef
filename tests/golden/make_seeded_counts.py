"""Writes seeded_8x8_seed2.json: the 256 seeded 8x8 boards of the benchmark
workload (4 blocked cells, region areas multiples of 5, mt19937_64 seed 2)
with their tiling counts by the oracle (12 free pentominoes x1, D4).

The oracle is pinned by tests/test_oracle.py; this fixture lets the GPU
tests check all 256 boards without re-running the CPU search."""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "d4-pieces_amd"))

from conftest import Oracle  # noqa: E402
import workloads  # noqa: E402


def main(n=256, seed=2, threads=8):
    oracle = Oracle(os.path.join(ROOT, "oracle", "liboracle.so"))
    p5 = oracle.enum_free(5)
    boards = workloads.seeded_8x8(n, seed=seed)
    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        counts = list(ex.map(lambda b: oracle.count(p5, [1] * 12, b), boards))
    json.dump({"seed": seed, "pieces": "12 free pentominoes x1, D4",
               "boards": boards, "counts": counts,
               "oracle_seconds": round(time.time() - t0, 1)},
              open(os.path.join(HERE, f"seeded_8x8_seed{seed}.json"), "w"))


if __name__ == "__main__":
    main()
