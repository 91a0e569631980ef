"""The host-only naming / ledger / polyomino code (d4_ledger.cpp) under
AddressSanitizer and UndefinedBehaviorSanitizer: tests/host/ledger_sanitize.cpp
is built with g++ against that one source file and run; any sanitizer report
or failed check fails the test.  (Device code is not built here: sanitizers
apply to host code only.)"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HOST = os.path.join(ROOT, "tests", "host")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_ledger_naming_polyominoes_sanitized(tmp_path):
    exe = str(tmp_path / "ledger_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "d4-pieces_amd", "csrc"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(ROOT, "d4-pieces_amd", "csrc", "d4_ledger.cpp"),
           os.path.join(HOST, "ledger_sanitize.cpp"),
           os.path.join(HOST, "set_error_stub.cpp"),
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env,
                       cwd=str(tmp_path))
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "fails 0" in r.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    assert "polyominoes n=8 sym=0xff: 369" in r.stdout
    assert "polyominoes n=8 sym=0x69: 704" in r.stdout
    assert "polyominoes n=8 sym=0x1: 2725" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_row_state_guards(tmp_path):
    """The 2^32 unique-state guard before compaction and the 64-bit
    compaction check (tests/host/guard_check.cpp), built host-only."""
    exe = str(tmp_path / "guard_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "d4-pieces_amd", "csrc"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(HOST, "guard_check.cpp"), os.path.join(HOST, "set_error_stub.cpp"),
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fails 0" in r.stdout
