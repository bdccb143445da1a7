"""The benchmark workload at its full size: R-MAT scale 20, edge factor 16,
seed 1, A * A in row blocks of at most 2^31 products (its 9.7e9 output
nonzeros exceed one uint32 DCSR).  Every row of the product is checked:

* the oracle's per-row digests (column count and an order-independent 64-bit
  hash of the columns, oracle.spgemm_digest: Gustavson without materialising
  the product) of all 546,835 stored rows of A equal the digests of the GPU
  result read back in chunks, for the plain two-pass and the learned one-pass
  modes; the GPU rows are exactly A's rows with a nonempty product and every
  row's columns are strictly ascending;
* row pointers and row ids of the two modes agree in full, the blocks'
  products add up to the workload total, and the product does not fit one
  uint32 DCSR.
"""
import numpy as np
import pytest

import bench
import clbool_amd as cb
import oracle
from clbool_amd import generators as gen

pytestmark = pytest.mark.gpu

CHUNK = 1 << 28  # columns read back per chunk (1 GiB)


@pytest.fixture(scope="module")
def rmat20(controls):
    r, c = gen.rmat(20, 16, 1)
    rpt, rows, cols = gen.to_dcsr(r, c)
    n = 1 << 20
    A = oracle.Dcsr(rpt, rows, cols, n, n)
    a = cb.matrix_dcsr_from_cpu(controls, rpt, rows, cols, n, n)
    prod = bench.row_products(rpt, rows, cols, n, rpt, rows)
    return A, a, prod, bench.row_blocks(prod, 1 << 31)


def _gpu_digests(ctl, m):
    """(rpt, rows, per-row digests, rows with unsorted columns) of a DCSR
    block on the device, its columns read back CHUNK at a time."""
    rpt = m.rpt_gpu.read(ctl, m.nzr + 1).astype(np.int64)
    rows = m.rows_gpu.read(ctl, m.nzr)
    h = np.zeros(m.nzr, np.uint64)
    bad = 0
    r = 0
    while r < m.nzr:
        start = int(rpt[r])
        r2 = int(np.searchsorted(rpt, start + CHUNK, side="right")) - 1
        r2 = min(max(r2, r + 1), m.nzr)
        cols = m.cols_gpu.read(ctl, int(rpt[r2]) - start, offset=4 * start)
        hh, b = oracle.dcsr_row_digest(rpt.astype(np.uint32), cols, base=start, r0=r, r1=r2)
        h[r:r2] = hh
        bad += b
        r = r2
    return rpt, rows, h, bad


def test_rmat20_every_row_matches_oracle(controls, rmat20):
    A, a, prod, blocks = rmat20
    bench.train_predictor(cb, gen, controls, lambda *x: None)
    total_products, total_nnz, checked_rows = 0, 0, 0
    for lo, hi in blocks:
        cnt, href = oracle.spgemm_digest(A, A, lo, hi)
        nz = cnt > 0
        want_rows = A.rows[lo:hi][nz]
        want_len = cnt[nz].astype(np.int64)
        want_h = href[nz]
        got = {}
        for mode in (cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED):
            M, st = cb.dcsr.spgemm(controls, a, a, mode=mode, row_begin=lo, row_end=hi)
            assert st["products"] == int(prod[lo:hi].sum())
            rpt, rows, h, bad = _gpu_digests(controls, M)
            what = f"block [{lo}, {hi}) mode {mode}"
            assert bad == 0, f"{what}: {bad} rows with unsorted or repeated columns"
            assert np.array_equal(rows, want_rows), f"{what}: row ids differ"
            assert np.array_equal(np.diff(rpt), want_len), f"{what}: row sizes differ"
            diff = np.flatnonzero(h != want_h)
            assert diff.size == 0, f"{what}: {diff.size} rows differ, first row {rows[diff[0]]}"
            got[mode] = (rpt, rows, M.nnz)
            del M
        assert np.array_equal(got[cb.SPGEMM_PLAIN][0], got[cb.SPGEMM_LEARNED][0])
        total_products += int(prod[lo:hi].sum())
        total_nnz += got[cb.SPGEMM_PLAIN][2]
        checked_rows += int(nz.sum())
    assert total_products == int(prod.sum())
    assert checked_rows > 500_000
    assert total_nnz > (1 << 32)  # the product does not fit one uint32 DCSR
