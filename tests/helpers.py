"""Shared helpers of the GPU tests: upload oracle matrices, compare results."""
import numpy as np

import clbool_amd as cb
import oracle


def upload(ctl, m: "oracle.Dcsr", nrows, ncols):
    return cb.matrix_dcsr_from_cpu(ctl, m.rpt, m.rows, m.cols, nrows, ncols)


def pairs_dcsr(ctl, rows, cols, nrows, ncols):
    m = oracle.pairs_to_dcsr(rows, cols, nrows, ncols)
    return m, upload(ctl, m, nrows, ncols)


def assert_dcsr_equal(ctl, got: "cb.matrix_dcsr", ref: "oracle.Dcsr", what=""):
    """compare_matrices (src/common/utils.cpp:42-56): nnz, then rpt, rows, cols."""
    assert got.nnz == ref.nnz, f"{what}: nnz {got.nnz} != {ref.nnz}"
    if ref.nnz == 0:
        return
    rpt, rows, cols = got.to_host(ctl)
    assert np.array_equal(rpt, ref.rpt), f"{what}: rpt differs"
    assert np.array_equal(rows, ref.rows), f"{what}: rows differ"
    if not np.array_equal(cols, ref.cols):
        d = np.flatnonzero(cols != ref.cols)[:5]
        raise AssertionError(f"{what}: cols differ at {d}: {cols[d]} vs {ref.cols[d]}")
