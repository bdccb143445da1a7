"""The oracle's CPU restatements under AddressSanitizer and
UndefinedBehaviorSanitizer: the oracle test suite runs against a sanitizer
build of oracle/oracle.c (the Makefile's `sanitize` target) in a child
process, and any out-of-bounds access, use after free or undefined
behaviour fails it.  Host code only."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    cc = shutil.which("gcc")
    if not cc:
        return None
    p = subprocess.run([cc, f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if p and os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_suite_under_asan_and_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc sanitizer runtimes not available")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ)
    env.update({
        # the sanitizer runtimes first, whatever the environment preloads after them
        "LD_PRELOAD": ":".join([asan, ubsan] + ([os.environ["LD_PRELOAD"]] if os.environ.get("LD_PRELOAD") else [])),
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "CLBOOL_ORACLE_LIB": os.path.join(ROOT, "oracle", "_asan", "liboracle.so"),
        "OMP_NUM_THREADS": "2",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle.py")], capture_output=True, text=True, env=env,
                       cwd=ROOT, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "passed" in r.stdout
