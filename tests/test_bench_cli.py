"""bench.py's command-line contract that holds without a GPU: a request for more
than one GPU exits at once with a message (this measurement is one GPU's), and the
defaults name the workload's launch shapes."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert not r.stdout.strip().startswith("{")  # no JSON line for N > 1


def test_bench_help_lists_launch_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--launch", "--serve-launch"):
        assert opt in r.stdout
