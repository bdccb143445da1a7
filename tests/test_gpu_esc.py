"""Expand-sort-compress: the dcsr::matrix_multiplication SpGEMM
(dcsr_matrix_multiplication.cpp:26-168) and the LDS ESC bin of the hash path,
against the oracle (matrix_multiplication_cpu, coo_utils.cpp:470-514).

* clb_dcsr_spgemm_esc on the reference's test sizes (test_operations.cpp
  multiplication cases), on power-law products whose long rows take the
  HBM sort path, on a hypersparse and a wide product, and on the edge cases
  (empty operands, no products, dimension mismatch).
* The hash path with esc_max_products set: rows bound for multi-window
  bitmaps go to the LDS ESC kernels, in the symbolic, numeric and one-pass
  modes, with identical results.
"""
import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import autotune as at
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,k", [(100, 3), (1024, 10), (3000, 50), (3800, 55)])
def test_esc_spgemm_reference_sizes(controls, size, k):
    # test_multiplication (test/src/operations.cpp:43-81 pattern): A = A * A
    r, c = oracle.generate_coo_pairs(size * k, size)
    A, a = pairs_dcsr(controls, r, c, size, size)
    got = cb.dcsr.matrix_multiplication(controls, a, a)
    assert_dcsr_equal(controls, got, oracle.spgemm_reference(A, A), f"esc {size}x{k}")
    ra, ca = gen.uniform_pairs(10000, 1024, 1024, size)
    rb, cb_ = gen.uniform_pairs(10000, 1024, 1024, size + 1)
    A, a = pairs_dcsr(controls, ra, ca, 1024, 1024)
    B, b = pairs_dcsr(controls, rb, cb_, 1024, 1024)
    got = cb.dcsr.matrix_multiplication(controls, a, b)
    assert_dcsr_equal(controls, got, oracle.spgemm_reference(A, B), "esc 1024^2 10k pairs")


@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 1), (12, 16, 2), (14, 8, 3)])
def test_esc_spgemm_power_law(controls, scale, ef, seed):
    n = 1 << scale
    A, a = pairs_dcsr(controls, *gen.rmat(scale, ef, seed), n, n)
    ref = oracle.spgemm(A, A)
    prod = oracle.workload(A, A)
    assert prod.max() > 8192  # some rows take the HBM sort path
    got = cb.dcsr.matrix_multiplication(controls, a, a)
    assert_dcsr_equal(controls, got, ref, f"esc rmat {scale}")
    assert controls.last_stats()["products"] == int(prod.sum())


def test_esc_spgemm_hypersparse_and_wide(controls):
    r, c = gen.hypersparse_zipf(seed=4, log_n=22, row_frac=0.02)
    n = 1 << 22
    A, a = pairs_dcsr(controls, r, c, n, n)
    assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication(controls, a, a), oracle.spgemm(A, A), "esc hyper")
    n2, wide = 1 << 12, 1 << 24
    r, c = gen.rmat(12, 16, 5)
    A2, a2 = pairs_dcsr(controls, r, c, n2, n2)
    cw = (c.astype(np.uint64) * 2654435761 % wide).astype(np.uint32)
    key = np.unique((r.astype(np.uint64) << np.uint64(32)) | cw)
    B2, b2 = pairs_dcsr(controls, (key >> np.uint64(32)).astype(np.uint32),
                        (key & np.uint64(0xFFFFFFFF)).astype(np.uint32), n2, wide)
    got = cb.dcsr.matrix_multiplication(controls, a2, b2)
    assert_dcsr_equal(controls, got, oracle.spgemm(A2, B2), "esc wide")
    # the hash path agrees
    assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication_hash(controls, a2, b2), oracle.spgemm(A2, B2), "hash")


def test_esc_spgemm_edge_cases(controls):
    r, c = gen.uniform_pairs(500, 100, 80, 3)
    A, a = pairs_dcsr(controls, r, c, 100, 80)
    out = cb.dcsr.matrix_multiplication(controls, a, cb.matrix_dcsr(80, 60))
    assert out.nnz == 0 and (out.nrows, out.ncols) == (100, 60)
    _, b = pairs_dcsr(controls, [79], [3], 80, 60)
    _, a2 = pairs_dcsr(controls, [0, 5], [1, 2], 100, 80)
    out = cb.dcsr.matrix_multiplication(controls, a2, b)
    assert out.nnz == 0 and (out.nrows, out.ncols) == (100, 60)
    with pytest.raises(cb.ClboolError):
        cb.dcsr.matrix_multiplication(controls, a, a)
    # README example (test/test_operations.cpp:284-291)
    A = oracle.pairs_to_dcsr([0, 0, 0, 2, 2, 4], [0, 1, 4, 2, 3, 2], 5, 5)
    B = oracle.pairs_to_dcsr([1, 1, 2, 3, 3, 3, 4], [0, 4, 2, 2, 3, 4, 2], 5, 5)
    a = cb.matrix_dcsr_from_cpu(controls, A.rpt, A.rows, A.cols, 5)
    b = cb.matrix_dcsr_from_cpu(controls, B.rpt, B.rows, B.cols, 5)
    assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication(controls, a, b), oracle.spgemm(A, B), "readme")


@pytest.mark.parametrize("mode", [cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED, cb.SPGEMM_BOUND])
def test_esc_bin_of_the_hash_path(controls, mode):
    # 2^22 columns: bitmap rows need several windows, so with the knob set
    # rows of up to 8192 products bound for a bitmap are sorted in LDS
    n, wide = 1 << 13, 1 << 22
    r, c = gen.rmat(13, 16, 71)
    A, a = pairs_dcsr(controls, r, c, n, n)
    cw = (c.astype(np.uint64) * 2654435761 % wide).astype(np.uint32)
    key = np.unique((r.astype(np.uint64) << np.uint64(32)) | cw)
    B, b = pairs_dcsr(controls, (key >> np.uint64(32)).astype(np.uint32),
                      (key & np.uint64(0xFFFFFFFF)).astype(np.uint32), n, wide)
    ref = oracle.spgemm(A, B)
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    p[-1] = 1.0
    cb.predictor_set_params(controls, p)
    base = at.tuning_dict(controls.tuning())
    try:
        at.apply(controls, {**base, "esc_max_products": 8192, "hash_max_slots": 1024})
        got, st = cb.dcsr.spgemm(controls, a, b, mode=mode, timing=True)
        assert_dcsr_equal(controls, got, ref, f"esc bin, mode {mode}")
        assert sum(st["bin_rows"][10:13]) > 0, st["bin_rows"]
    finally:
        at.apply(controls, base)
