"""clb_prefix_sum_noalloc through the C-ABI: the same results as
clb_prefix_sum (and the oracle's scan) on two calls in a row, the second on a
smaller input, and its size limit."""
import ctypes as C

import numpy as np
import pytest

import clbool_amd as cb
import oracle

pytestmark = pytest.mark.gpu


def _noalloc(ctl, buf, n):
    fn = cb.lib().clb_prefix_sum_noalloc
    fn.restype = C.c_int32
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
    t = C.c_uint32()
    rc = fn(ctl.handle, buf.handle, n, C.byref(t))
    return rc, t.value


def test_noalloc_prefix_sum_matches_prefix_sum(controls):
    rng = np.random.default_rng(12)
    for n in (5_000_001, 70_001, 4096, 1):  # each call smaller than the one before
        x = rng.integers(0, 700, n, dtype=np.uint32)
        a = cb.Buffer.from_array(controls, x)
        b = cb.Buffer.from_array(controls, x)
        ta = cb.prefix_sum(controls, a, n)
        rc, tb = _noalloc(controls, b, n)
        assert rc == 0, cb.lib().clb_last_error()
        ga, gb = a.read(controls, n), b.read(controls, n)
        ref, rt = oracle.exclusive_scan(x)
        assert ta == tb == rt
        assert np.array_equal(ga, gb) and np.array_equal(gb, ref), f"n={n}"


def test_noalloc_prefix_sum_limits(controls):
    x = np.arange(16, dtype=np.uint32)
    b = cb.Buffer.from_array(controls, x)
    rc, t = _noalloc(controls, b, (1 << 28) + 1)   # above the scratch's limit: refused before any launch
    assert rc == 1 + 11  # CLBOOL_INVALID_ARGUMENT
    rc, t = _noalloc(controls, b, 17)              # larger than the buffer
    assert rc == 1 + 11
    rc, t = _noalloc(controls, b, 0)
    assert rc == 0 and t == 0
