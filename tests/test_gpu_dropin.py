"""The reference's test flows written in C++ against include/clbool.hpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_dropin_program():
    exe = os.path.join(HERE, "cpp", "test_dropin")
    if not os.path.exists(exe):
        subprocess.run(["sh", os.path.join(HERE, "cpp", "build.sh")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "0 failed" in r.stdout
