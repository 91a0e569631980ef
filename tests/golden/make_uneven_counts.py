"""Writes uneven_4096_seed3.json: oracle tiling counts of the 4096 uneven
puzzles (workloads.uneven_instances(4096, seed=3)) that the oracle's
backtracker finishes within a budget of DFS nodes, plus a digest of the whole
instance list and its first instances written out in full.

Run from the repo root after building the oracle (make -C oracle):
    python tests/golden/make_uneven_counts.py [--budget 10000000] [--jobs 8]
"""
import argparse
import hashlib
import json
import os
import sys
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "d4-pieces_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

_inst = None
_oracle = None


def _init():
    global _inst, _oracle
    import workloads
    from conftest import Oracle
    _inst = workloads.uneven_instances(4096, seed=3)
    _oracle = Oracle(os.path.join(ROOT, "oracle", "liboracle.so"))


def _solve(args):
    i, budget = args
    shapes, counts, rows, _ = _inst[i]
    n, done = _oracle.count_budget(shapes, counts, rows, budget)
    return i, n, done


def digest(instances):
    h = hashlib.sha256()
    for shapes, counts, rows, w in instances:
        h.update(json.dumps([shapes, counts, rows, w]).encode())
    return h.hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget", type=int, default=10_000_000)
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    _init()
    out = {"generator": "workloads.uneven_instances(4096, seed=3)",
           "sha256": digest(_inst),
           "first": [list(x) for x in _inst[:4]],
           "budget_nodes": args.budget, "counts": {}}
    with Pool(args.jobs, initializer=_init) as p:
        for k, (i, n, done) in enumerate(
                p.imap_unordered(_solve, [(i, args.budget) for i in range(len(_inst))],
                                 chunksize=4)):
            if done:
                out["counts"][str(i)] = n
            if k % 256 == 0:
                print(k, len(out["counts"]), flush=True)
    out["counts"] = dict(sorted(out["counts"].items(), key=lambda kv: int(kv[0])))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uneven_4096_seed3.json")
    json.dump(out, open(path, "w"), indent=0)
    print("pinned", len(out["counts"]), "of", len(_inst))


if __name__ == "__main__":
    main()
