"""The autotuner on the GPU: every knob setting it accepts reproduces the
oracle's product exactly, and the device-side digest it compares by equals a
direct evaluation of its definition and changes when one entry changes."""
import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import autotune as at
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1


def _mix64(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _digest_py(m, nrows, ncols):
    h = 0
    for i in range(m.nzr):
        for q in range(int(m.rpt[i]), int(m.rpt[i + 1])):
            h = (h + _mix64(_mix64(q) ^ ((int(m.rows[i]) << 32) | int(m.cols[q])))) & M64
    return h ^ ((((m.nnz << 32) | m.nzr) * 0x9E3779B97F4A7C15) & M64) ^ ((nrows << 32) | ncols)


def test_digest_matches_definition_and_detects_changes(controls):
    r, c = gen.uniform_pairs(3000, 300, 200, 8)
    A, a = pairs_dcsr(controls, r, c, 300, 200)
    assert cb.dcsr.digest(controls, a) == _digest_py(A, 300, 200)
    # one column changed (the digest does not need a canonical matrix)
    cols = A.cols.copy()
    cols[1500] = (cols[1500] + 1) % 200
    b = cb.matrix_dcsr_from_cpu(controls, A.rpt, A.rows, cols, 300, 200)
    assert cb.dcsr.digest(controls, b) != cb.dcsr.digest(controls, a)
    # two columns swapped between positions
    cols = A.cols.copy()
    cols[[10, 2000]] = cols[[2000, 10]]
    b1 = cb.matrix_dcsr_from_cpu(controls, A.rpt, A.rows, cols, 300, 200)
    assert A.cols[10] == A.cols[2000] or cb.dcsr.digest(controls, b1) != cb.dcsr.digest(controls, a)
    # a row id moved to another row
    rows = A.rows.copy()
    rows[-1] = 299 if rows[-1] != 299 else 298
    b2 = cb.matrix_dcsr_from_cpu(controls, A.rpt, rows, A.cols, 300, 200)
    assert cb.dcsr.digest(controls, b2) != cb.dcsr.digest(controls, a)
    empty = cb.matrix_dcsr(300, 200)
    assert cb.dcsr.digest(controls, empty) == _digest_py(oracle.Dcsr(np.zeros(1, np.uint32), np.zeros(0, np.uint32),
                                                                      np.zeros(0, np.uint32), 300, 200), 300, 200)


@pytest.mark.parametrize("mode", [cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED])
def test_accepted_knobs_reproduce_the_oracle(controls, mode):
    n = 1 << 13
    A, a = pairs_dcsr(controls, *gen.rmat(13, 16, 61), n, n)
    ref = oracle.spgemm(A, A)
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    p[-1] = 10.0 / 32.0
    cb.predictor_set_params(controls, p)
    base = at.tuning_dict(controls.tuning())
    blocks = [(0, A.nzr // 3), (A.nzr // 3, A.nzr)]
    try:
        knobs = ["hash_max_slots", "bitmap_min_products", "tiny_max_products", "gather_combine"]
        if mode == cb.SPGEMM_LEARNED:
            knobs += ["bitmap_pred_caps", "learned_margin_pct"]
        res = at.autotune(controls, a, a, blocks, mode=mode, repeats=2, knobs=knobs, passes=1, min_gain=-1.0)
        # min_gain -1 accepts every setting that keeps the result: the final
        # knobs are the last value of each knob, all checked by digest
        at.apply(controls, res["tuning"])
        got, _ = cb.dcsr.spgemm(controls, a, a, mode=mode)
        assert_dcsr_equal(controls, got, ref, f"autotuned {res['tuning']}")
        assert res["evaluated"] > len(knobs)
    finally:
        at.apply(controls, base)
