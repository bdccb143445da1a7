"""Host-side logic of the benchmark: per-row product counts, the row blocks
that keep every block's result within uint32 indices, and the one-GPU guard.
No GPU calls."""
import subprocess
import sys
import os

import numpy as np

import bench
import oracle
from clbool_amd import generators as gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _small_rmat(scale=10, ef=8, seed=3):
    r, c = gen.rmat(scale, ef, seed)
    return gen.to_dcsr(r, c), 1 << scale


def test_row_products_match_brute_force():
    (rpt, rows, cols), n = _small_rmat()
    prod = bench.row_products(rpt, rows, cols, n, rpt, rows)
    blen = {int(r): int(rpt[i + 1] - rpt[i]) for i, r in enumerate(rows)}
    want = [sum(blen.get(int(k), 0) for k in cols[rpt[i]:rpt[i + 1]]) for i in range(rows.size)]
    assert prod.tolist() == want
    # the products bound each row's exact output size (Gustavson on the oracle)
    A = oracle.Dcsr(rpt, rows, cols, n, n)
    C = oracle.spgemm(A, A)
    size = dict(zip(C.rows.tolist(), np.diff(C.rpt.astype(np.int64)).tolist()))
    assert all(size.get(int(r), 0) <= min(int(p), n) for r, p in zip(rows, prod))


def test_row_blocks_cover_rows_and_respect_budget():
    rng = np.random.default_rng(5)
    prod = rng.integers(0, 5000, 3000)
    prod[17] = 40000  # a row above the budget gets a block of its own
    for budget, ncols in ((10000, None), (10000, 3000), (1 << 40, None)):
        blocks = bench.row_blocks(prod, budget, ncols=ncols)
        w = np.minimum(prod, ncols) if ncols else prod
        assert blocks[0][0] == 0 and blocks[-1][1] == prod.size
        assert all(b[1] == nb[0] for b, nb in zip(blocks, blocks[1:]))
        for lo, hi in blocks:
            assert hi > lo
            assert w[lo:hi].sum() <= budget or hi - lo == 1
    assert len(bench.row_blocks(prod, 1 << 40)) == 1


def test_exact_size_blocks_are_never_more_than_bound_blocks():
    """Blocks cut by exact output sizes never outnumber the ones cut by the
    output bound sum(min(products, ncols)) at the same budget."""
    (rpt, rows, cols), n = _small_rmat(11, 16, 2)
    prod = bench.row_products(rpt, rows, cols, n, rpt, rows)
    A = oracle.Dcsr(rpt, rows, cols, n, n)
    C = oracle.spgemm(A, A)
    sizes = np.zeros(rows.size, np.int64)
    sizes[np.searchsorted(rows, C.rows)] = np.diff(C.rpt.astype(np.int64))
    budget = int(np.minimum(prod, n).sum() // 7)
    bound = bench.row_blocks(prod, budget, ncols=n)
    exact = bench.row_blocks(sizes, budget)
    assert len(exact) <= len(bound)
    assert all(sizes[lo:hi].sum() <= budget for lo, hi in exact)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
