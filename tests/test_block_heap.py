"""The device pool's block bookkeeping (clbool_amd/csrc/block_heap.hpp) on
the host: tests/cpp/test_block_heap.cpp runs randomised take/release
sequences against it (containment, no overlap, byte accounting, coalescing,
whole-segment collection), built with AddressSanitizer and
UndefinedBehaviorSanitizer where the compiler has them."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_block_heap(tmp_path):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    src = os.path.join(ROOT, "tests", "cpp", "test_block_heap.cpp")
    exe = str(tmp_path / "test_block_heap")
    base = [cxx, "-std=c++17", "-O1", "-g", "-Wall", "-Werror", src, "-o", exe]
    r = subprocess.run(base[:1] + ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"] + base[1:],
                       capture_output=True, text=True)
    if r.returncode != 0:  # no sanitizer runtimes: plain build
        r = subprocess.run(base, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "block heap: ok" in r.stdout
