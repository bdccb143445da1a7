"""The benchmark workload at its full size: R-MAT scale 20, edge factor 16,
seed 1, A * A in row blocks of at most 2^31 products (its 9.7e9 output
nonzeros exceed one uint32 DCSR).  The oracle cannot redo the whole product
in a test, so the full-size checks are size-independent properties:

* plain two-pass and learned one-pass results agree: row pointers and row ids
  in full, columns in three 2^20-wide windows of every block;
* every row's size is at most min(products of the row, ncols), rows are
  strictly ascending, and the blocks' products add up to the workload total;
* 48 seeded rows per block equal the oracle's Gustavson product exactly.
"""
import numpy as np
import pytest

import bench
import clbool_amd as cb
import oracle
from clbool_amd import generators as gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat20(controls):
    r, c = gen.rmat(20, 16, 1)
    rpt, rows, cols = gen.to_dcsr(r, c)
    n = 1 << 20
    A = oracle.Dcsr(rpt, rows, cols, n, n)
    a = cb.matrix_dcsr_from_cpu(controls, rpt, rows, cols, n, n)
    prod = bench.row_products(rpt, rows, cols, n, rpt, rows)
    return A, a, prod, bench.row_blocks(prod, 1 << 31)


def _window_cols(ctl, m, k=1 << 20):
    k = min(m.nnz, k)
    return [m.cols_gpu.read(ctl, k, offset=4 * off) for off in (0, (m.nnz - k) // 2, m.nnz - k)]


def test_rmat20_full_product_properties(controls, rmat20):
    A, a, prod, blocks = rmat20
    bench.train_predictor(cb, gen, controls, lambda *x: None)
    rng = np.random.default_rng(7)
    total_products, total_nnz = 0, 0
    for lo, hi in blocks:
        P, st = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_PLAIN, row_begin=lo, row_end=hi)
        L, sl = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED, row_begin=lo, row_end=hi)
        assert st["products"] == sl["products"] == int(prod[lo:hi].sum())
        total_products += st["products"]
        assert (P.nnz, P.nzr) == (L.nnz, L.nzr), f"block {lo}: plain and learned sizes differ"
        prpt, prows = P.rpt_gpu.read(controls, P.nzr + 1), P.rows_gpu.read(controls, P.nzr)
        assert np.array_equal(prpt, L.rpt_gpu.read(controls, L.nzr + 1))
        assert np.array_equal(prows, L.rows_gpu.read(controls, L.nzr))
        for x, y in zip(_window_cols(controls, P), _window_cols(controls, L)):
            assert np.array_equal(x, y), f"block {lo}: column windows differ"
        # structure: ascending rows, each of them a row of the block, sizes
        # bounded by the row's products and the width
        assert np.all(np.diff(prows.astype(np.int64)) > 0)
        pos = np.searchsorted(A.rows, prows)
        assert np.all((pos >= lo) & (pos < hi)) and np.array_equal(A.rows[pos], prows)
        lens = np.diff(prpt.astype(np.int64))
        assert np.all(lens >= 1) and np.all(lens <= np.minimum(prod[pos], 1 << 20))
        # exact rows against the oracle
        sel = np.sort(rng.choice(np.arange(lo, hi), size=min(48, hi - lo), replace=False)).astype(np.uint32)
        ref = oracle.spgemm(A, A, row_sel=sel)
        for q in range(ref.nzr):
            i = int(np.searchsorted(prows, ref.rows[q]))
            assert i < P.nzr and prows[i] == ref.rows[q]
            got = P.cols_gpu.read(controls, int(prpt[i + 1] - prpt[i]), offset=4 * int(prpt[i]))
            assert np.array_equal(got, ref.cols[ref.rpt[q]:ref.rpt[q + 1]]), f"row {ref.rows[q]}"
        total_nnz += P.nnz
        del P, L
    assert total_products == int(prod.sum())
    assert total_nnz > (1 << 32)  # the product does not fit one uint32 DCSR
