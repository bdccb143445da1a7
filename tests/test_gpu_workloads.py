"""The closure and RPQ workloads at their full sizes, against the oracle.

* Closure: 64 label matrices on 2^18 vertices (Zipf(1.1) sizes over
  2^10..2^20, R-MAT structure), 32 seeded triples X |= Y * Z, two steps.
  After each step the per-triple plain mode (the reference's call pattern:
  one matrix_multiplication_hash and one addition per triple) and the
  batched learned mode hold identical matrices, every X's stored rows are its
  old rows plus the rows its products reach, and 64 sampled rows of every X
  equal old row | union of the oracle's product rows (coo_utils.cpp:316-324
  for the union, :470-514 for the products).
* RPQ input: the Kronecker product of a 64 x 64 automaton (200 uniform
  nonzeros) with an R-MAT scale-22, edge-factor-32 graph, in its 60 blocks of
  one automaton row (the whole product has 2.6e10 nonzeros, over uint32):
  per block nnz = |M_i| * nnz(G), row ids and row pointers in closed form,
  and sampled rows equal the oracle's kronecker_product_cpu
  (coo_utils.cpp:327-342).
* One closure step K * K of the same automaton with the graph at scale 16
  (K: 2^22 rows), learned one pass in planned row blocks: the blocks'
  products add up to products(M * M) * products(G * G) row by row, and 48
  sampled rows of every block equal the oracle's product.
"""
import numpy as np
import pytest

import bench
import clbool_amd as cb
import oracle
from clbool_amd import generators as gen

pytestmark = pytest.mark.gpu


def _host(ctl, m, n):
    rpt, rows, cols = m.to_host(ctl) if m.nnz else (np.zeros(1, np.uint32), np.zeros(0, np.uint32),
                                                     np.zeros(0, np.uint32))
    return oracle.Dcsr(rpt.astype(np.uint32), rows.astype(np.uint32), cols.astype(np.uint32), n, n)


def _row(m, r):
    i = int(np.searchsorted(m.rows, r))
    if i < m.nzr and m.rows[i] == r:
        return m.cols[m.rpt[i]:m.rpt[i + 1]]
    return np.zeros(0, np.uint32)


def _product_rows(Y, Z, n):
    """Stored rows of Y * Z: rows of Y with a column that is a nonempty row of Z."""
    zlen = np.zeros(n, np.int64)
    zlen[Z.rows] = np.diff(Z.rpt.astype(np.int64))
    if Y.nnz == 0:
        return np.zeros(0, np.uint32)
    hit = (zlen[Y.cols] > 0).astype(np.int64)
    per_row = np.add.reduceat(hit, Y.rpt[:-1].astype(np.int64))
    return Y.rows[per_row > 0]


def test_c4_closure_steps_at_workload_size(controls):
    n = 1 << 18
    labs = gen.closure_labels(n_labels=64, log_n=18, lo=1 << 10, hi=1 << 20, seed=0)
    triples = gen.closure_triples(64, 32, seed=1)
    host = [oracle.pairs_to_dcsr(r, c, n, n) for r, c in labs]
    bench.train_predictor(cb, gen, controls, lambda *x: None)
    per = [cb.matrix_dcsr_from_cpu(controls, m.rpt, m.rows, m.cols, n) for m in host]
    bat = [cb.matrix_dcsr_from_cpu(controls, m.rpt, m.rows, m.cols, n) for m in host]
    rng = np.random.default_rng(17)
    for step in range(2):
        sa = cb.closure_step(controls, per, triples, mode=cb.CLOSURE_PER_TRIPLE, spgemm_mode=cb.SPGEMM_PLAIN)
        sb = cb.closure_step(controls, bat, triples, mode=cb.CLOSURE_BATCHED, spgemm_mode=cb.SPGEMM_LEARNED)
        assert sa["products"] == sb["products"] and sa["changed"] == sb["changed"] > 0
        new = [_host(controls, m, n) for m in per]
        for k in range(64):
            b = _host(controls, bat[k], n)
            assert (new[k].nnz, new[k].nzr) == (b.nnz, b.nzr), f"step {step} label {k}: modes differ in size"
            assert np.array_equal(new[k].rpt, b.rpt) and np.array_equal(new[k].rows, b.rows)
            assert np.array_equal(new[k].cols, b.cols), f"step {step} label {k}: modes differ"
        for x in sorted({t[0] for t in triples}):
            srcs = [(y, z) for tx, y, z in triples if tx == x]
            want_rows = np.unique(np.concatenate([host[x].rows] + [_product_rows(host[y], host[z], n)
                                                                    for y, z in srcs]))
            assert np.array_equal(new[x].rows, want_rows), f"step {step} X={x}: stored rows differ"
            for r in rng.choice(new[x].rows, size=min(64, new[x].nzr), replace=False):
                parts = [_row(host[x], r)]
                for y, z in srcs:
                    i = np.searchsorted(host[y].rows, r)
                    if i < host[y].nzr and host[y].rows[i] == r:
                        parts.append(oracle.spgemm(host[y], host[z], row_sel=np.array([i], np.uint32)).cols)
                want = np.unique(np.concatenate(parts))
                assert np.array_equal(_row(new[x], r), want), f"step {step} X={x} row {r}"
        host = new
    assert sum(m.nnz for m in host) > 4 * sum(oracle.pairs_to_dcsr(r, c, n, n).nnz for r, c in labs)


def test_c5_kronecker_blocks_at_workload_size(controls):
    mr, mc = gen.uniform_pairs(200, 64, 64, 1)
    M = oracle.pairs_to_dcsr(mr, mc, 64, 64)
    gr, gc = gen.rmat(22, 32, 1)
    n = 1 << 22
    G = oracle.pairs_to_dcsr(gr, gc, n, n)
    del gr, gc
    g = cb.matrix_dcsr_from_cpu(controls, G.rpt, G.rows, G.cols, n, n)
    glen = np.diff(G.rpt.astype(np.int64))
    rng = np.random.default_rng(23)
    total = 0
    for q in range(M.nzr):
        lo, hi = int(M.rpt[q]), int(M.rpt[q + 1])
        mrow = np.full(hi - lo, M.rows[q], np.uint32)
        m1 = cb.matrix_dcsr_from_cpu(controls, np.array([0, hi - lo], np.uint32), M.rows[q:q + 1],
                                     M.cols[lo:hi], 64, 64)
        K = cb.dcsr.kronecker_product(controls, m1, g)
        assert (K.nrows, K.ncols) == (64 * n, 64 * n)
        assert K.nnz == (hi - lo) * G.nnz and K.nzr == G.nzr, f"block {q}"
        rows = K.rows_gpu.read(controls, K.nzr)
        rpt = K.rpt_gpu.read(controls, K.nzr + 1)
        assert np.array_equal(rows, (int(M.rows[q]) * n + G.rows.astype(np.int64)).astype(np.uint32))
        assert np.array_equal(rpt.astype(np.int64), (hi - lo) * G.rpt.astype(np.int64))
        # sampled rows against kronecker_product_cpu of the automaton row and
        # the sampled graph rows
        sel = np.sort(rng.choice(G.nzr, 16, replace=False))
        sr = np.repeat(G.rows[sel], glen[sel])
        sc = np.concatenate([G.cols[G.rpt[i]:G.rpt[i + 1]] for i in sel])
        er, ec = oracle.kron(mrow, M.cols[lo:hi], sr, sc, n, n)
        ref = oracle.pairs_to_dcsr(er, ec, 64 * n, 64 * n)
        for j, i in enumerate(sel):
            got = K.cols_gpu.read(controls, int(rpt[i + 1] - rpt[i]), offset=4 * int(rpt[i]))
            assert ref.rows[j] == rows[i]
            assert np.array_equal(got, ref.cols[ref.rpt[j]:ref.rpt[j + 1]]), f"block {q} row {rows[i]}"
        total += K.nnz
        del K, m1
    assert total == M.nnz * G.nnz and total > (1 << 32)


def test_c5_closure_step_blocks(controls):
    mr, mc = gen.uniform_pairs(200, 64, 64, 1)
    M = oracle.pairs_to_dcsr(mr, mc, 64, 64)
    hr, hc = gen.rmat(16, 32, 1)
    m = 1 << 16
    H = oracle.pairs_to_dcsr(hr, hc, m, m)
    k = cb.dcsr.kronecker_product(controls, cb.matrix_dcsr_from_cpu(controls, M.rpt, M.rows, M.cols, 64, 64),
                                  cb.matrix_dcsr_from_cpu(controls, H.rpt, H.rows, H.cols, m, m))
    K = _host(controls, k, 64 * m)
    # K's rows (i, g) in closed form
    assert np.array_equal(K.rows, (np.repeat(M.rows.astype(np.int64) * m, H.nzr) +
                                   np.tile(H.rows.astype(np.int64), M.nzr)).astype(np.uint32))
    # products of row (i, g) of K * K: sum_{j in M_i} |M_j| * products(H_g)
    mlen = np.zeros(64, np.int64)
    mlen[M.rows] = np.diff(M.rpt.astype(np.int64))
    two = np.add.reduceat(mlen[M.cols], M.rpt[:-1].astype(np.int64))
    pg = bench.row_products(H.rpt, H.rows, H.cols, m, H.rpt, H.rows)
    prod = np.concatenate([int(s) * pg for s in two])
    bench.train_predictor(cb, gen, controls, lambda *x: None)
    plan = cb.dcsr.plan(controls, k, k, mode=cb.SPGEMM_LEARNED)
    assert plan.info()["products"] == int(prod.sum())
    rng = np.random.default_rng(29)
    blocks, nnz = 0, 0
    b = 0
    while b < plan.nblocks:
        C, st = plan.run(b)
        lo, hi = plan.block(b)
        assert st["products"] == int(prod[lo:hi].sum())
        crows = C.rows_gpu.read(controls, C.nzr)
        crpt = C.rpt_gpu.read(controls, C.nzr + 1)
        sel = np.sort(rng.choice(np.arange(lo, hi), size=min(48, hi - lo), replace=False)).astype(np.uint32)
        ref = oracle.spgemm(K, K, row_sel=sel)
        for q in range(ref.nzr):
            i = int(np.searchsorted(crows, ref.rows[q]))
            assert i < C.nzr and crows[i] == ref.rows[q]
            got = C.cols_gpu.read(controls, int(crpt[i + 1] - crpt[i]), offset=4 * int(crpt[i]))
            assert np.array_equal(got, ref.cols[ref.rpt[q]:ref.rpt[q + 1]]), f"block {b} row {ref.rows[q]}"
        blocks += 1
        nnz += C.nnz
        del C
        b += 1
    assert blocks > 10 and nnz > (1 << 36)
