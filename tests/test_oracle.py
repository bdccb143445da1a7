"""Pins the CPU oracle (CPU only).

Each oracle function is checked against something that does not restate its
own formula: the reference's own FastRandom (oracle/_ref and a fixture made
from it), hand-derived answers of the reference's test inputs, dense boolean
brute force, numpy set operations, finite differences.  A planted mistake
(a dropped column, a transposed operand) must fail the same checks.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def dense_from_dcsr(m: oracle.Dcsr, nrows, ncols):
    d = np.zeros((nrows, ncols), np.uint8)
    for i in range(m.nzr):
        d[m.rows[i], m.cols[m.rpt[i]:m.rpt[i + 1]]] = 1
    return d


def dcsr_from_dense(d):
    r, c = np.nonzero(d)
    return oracle.pairs_to_dcsr(r.astype(np.uint32), c.astype(np.uint32), d.shape[0], d.shape[1])


def random_dense(rng, m, n, density):
    return (rng.random((m, n)) < density).astype(np.uint8)


# ---- FastRandom ------------------------------------------------------------
def test_fastrandom_matches_fixture():
    for case in _load("fastrandom.json")["cases"]:
        got = oracle.fastrandom_stream(case["seed"], case["min"], case["max"], len(case["draws"]))
        assert got.tolist() == case["draws"], case["seed"]


def test_fastrandom_matches_reference_build():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (clBool tree absent)")
    assert oracle.ref_lib().ref_unsigned_long_bits() == 64
    for seed in (1, 10000, 150000, 190000, 2 ** 31 - 1, 3000 * 50):
        for lo, hi in ((0, 1023), (0, 2999), (0, 2 ** 31 - 1), (-100, 100)):
            a = oracle.fastrandom_stream(seed, lo, hi, 4096)
            b = oracle.ref_fastrandom_stream(seed, lo, hi, 4096)
            assert np.array_equal(a, b), (seed, lo, hi)


def _xorshf96_numpy(seed, n, hi):
    """Independent numpy restatement of xorshf96 (fast_random.h:22-32)."""
    x, y, z = np.uint64(seed), np.uint64(362436069), np.uint64(521288629)
    out = np.zeros(n, np.uint64)
    with np.errstate(over="ignore"):
        for i in range(n):
            x ^= x << np.uint64(16)
            x ^= x >> np.uint64(5)
            x ^= x << np.uint64(1)
            t = x
            x, y = y, z
            z = t ^ x ^ y
            out[i] = z % np.uint64(hi + 1)
    return out


def test_generate_coo_pairs_matches_fixture_and_numpy():
    for case in _load("generated_pairs.json")["cases"]:
        r, c = oracle.generate_coo_pairs(case["size"] * case["k"], case["size"])
        key = (r.astype(np.uint64) << np.uint64(32)) | c.astype(np.uint64)
        assert r.size == case["n"]
        assert int(r.astype(np.uint64).sum()) == case["sum_rows"]
        assert int(c.astype(np.uint64).sum()) == case["sum_cols"]
        assert int(np.bitwise_xor.reduce(key)) == case["xor_keys"]
        assert [[int(a), int(b)] for a, b in zip(r[:8], c[:8])] == case["first"]
    # small case against an independent numpy generator + np.unique
    n, size = 15 * 7, 15
    d = _xorshf96_numpy(n, 2 * n, size - 1)
    rows, cols = d[0::2], d[1::2]
    key = np.unique((rows << np.uint64(32)) | cols)
    r, c = oracle.generate_coo_pairs(n, size)
    assert np.array_equal(r, (key >> np.uint64(32)).astype(np.uint32))
    assert np.array_equal(c, (key & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_generate_zero_size():
    r, c = oracle.generate_coo_pairs(0, 0)
    assert r.size == 0
    with pytest.raises(ValueError):
        oracle.generate_coo_pairs(5, 0)


# ---- SpGEMM -----------------------------------------------------------------
def test_readme_known_answer():
    ka = _load("known_answers.json")["readme_product"]
    A = oracle.pairs_to_dcsr(ka["a"]["rows"], ka["a"]["cols"], 5, 5)
    B = oracle.pairs_to_dcsr(ka["b"]["rows"], ka["b"]["cols"], 5, 5)
    for f in (oracle.spgemm_reference, oracle.spgemm):
        C = f(A, B)
        assert C.rpt.tolist() == ka["c"]["rpt"]
        assert C.rows.tolist() == ka["c"]["rows"]
        assert C.cols.tolist() == ka["c"]["cols"]


@pytest.mark.parametrize("m,k,n,density", [(1, 1, 1, 1.0), (7, 5, 9, 0.3), (64, 64, 64, 0.05), (100, 37, 250, 0.1),
                                           (200, 200, 200, 0.02), (33, 300, 17, 0.5)])
def test_spgemm_oracles_match_dense_brute_force(m, k, n, density):
    rng = np.random.default_rng(m * 1000 + k + n)
    da, db = random_dense(rng, m, k, density), random_dense(rng, k, n, density)
    dc = oracle.dense_bool_mm(da, db)
    assert np.array_equal(dc, ((da.astype(np.int64) @ db.astype(np.int64)) > 0).astype(np.uint8))
    A, B = dcsr_from_dense(da), dcsr_from_dense(db)
    A.nrows, A.ncols, B.nrows, B.ncols = m, k, k, n
    for f in (oracle.spgemm_reference, oracle.spgemm):
        C = f(A, B)
        assert np.array_equal(dense_from_dcsr(C, m, n), dc)
        # canonical form: rows strictly ascending, columns sorted unique, no empty rows
        assert np.all(np.diff(C.rows.astype(np.int64)) > 0)
        assert np.all(np.diff(C.rpt.astype(np.int64)) > 0)
        for i in range(C.nzr):
            seg = C.cols[C.rpt[i]:C.rpt[i + 1]].astype(np.int64)
            assert np.all(np.diff(seg) > 0)


def test_planted_mistakes_are_caught():
    rng = np.random.default_rng(5)
    da, db = random_dense(rng, 40, 30, 0.15), random_dense(rng, 30, 50, 0.15)
    dc = oracle.dense_bool_mm(da, db)
    C = oracle.spgemm(dcsr_from_dense(da), dcsr_from_dense(db))
    dropped = oracle.Dcsr(C.rpt.copy(), C.rows.copy(), C.cols.copy())
    dropped.cols[0] = dropped.cols[1] if dropped.nnz > 1 else dropped.cols[0] + 1  # one column lost
    assert not np.array_equal(dense_from_dcsr(dropped, 40, 50), dc)
    # transposed operand gives a different product on a non-symmetric input
    assert not np.array_equal(oracle.dense_bool_mm(db.T.copy(), da.T.copy()).T, oracle.dense_bool_mm(da, da.T.copy()))


def test_identity_and_empty():
    rng = np.random.default_rng(1)
    da = random_dense(rng, 50, 50, 0.1)
    A = dcsr_from_dense(da)
    eye = dcsr_from_dense(np.eye(50, dtype=np.uint8))
    for f in (oracle.spgemm_reference, oracle.spgemm):
        C = f(eye, A)
        assert np.array_equal(C.rpt, A.rpt) and np.array_equal(C.rows, A.rows) and np.array_equal(C.cols, A.cols)
        C = f(A, eye)
        assert np.array_equal(C.cols, A.cols)
    Z = oracle.Dcsr(np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32), 50, 50)
    assert oracle.spgemm(A, Z).nnz == 0
    assert oracle.spgemm_reference(A, Z).nnz == 0


def test_reference_quirk_empty_b_row():
    # coo_utils.cpp:489-502: a found-but-empty B row still records the A row
    A = oracle.Dcsr(np.array([0, 1], np.uint32), np.array([3], np.uint32), np.array([2], np.uint32))
    B = oracle.Dcsr(np.array([0, 0], np.uint32), np.array([2], np.uint32), np.zeros(0, np.uint32))
    C = oracle.spgemm_reference(A, B)
    assert C.rows.tolist() == [3] and C.rpt.tolist() == [0, 0] and C.nnz == 0


def test_gustavson_row_selection_and_reference_agree():
    r, c = oracle.generate_coo_pairs(3000 * 5, 3000)
    A = oracle.pairs_to_dcsr(r, c, 3000, 3000)
    full = oracle.spgemm(A, A)
    ref = oracle.spgemm_reference(A, A)
    assert np.array_equal(full.rpt, ref.rpt) and np.array_equal(full.cols, ref.cols)
    sel = np.arange(0, A.nzr, 7, dtype=np.uint32)
    part = oracle.spgemm(A, A, row_sel=sel)
    fr = {int(full.rows[i]): full.cols[full.rpt[i]:full.rpt[i + 1]] for i in range(full.nzr)}
    for i in range(part.nzr):
        assert np.array_equal(part.cols[part.rpt[i]:part.rpt[i + 1]], fr[int(part.rows[i])])


def test_workload_brute_force():
    rng = np.random.default_rng(2)
    da, db = random_dense(rng, 60, 40, 0.2), random_dense(rng, 40, 70, 0.2)
    A, B = dcsr_from_dense(da), dcsr_from_dense(db)
    w = oracle.workload(A, B)
    expect = (da.astype(np.int64) @ db.sum(axis=1).astype(np.int64))[A.rows]
    assert np.array_equal(w.astype(np.int64), expect)


# ---- addition, Kronecker, scan, conversions ---------------------------------
def test_coo_add_is_set_union():
    rng = np.random.default_rng(3)
    da, db = random_dense(rng, 80, 90, 0.1), random_dense(rng, 80, 90, 0.1)
    ar, ac = np.nonzero(da)
    br, bc = np.nonzero(db)
    r, c = oracle.coo_add(ar, ac, br, bc)
    er, ec = np.nonzero(da | db)
    assert np.array_equal(r, er) and np.array_equal(c, ec)


def test_kron_matches_numpy():
    rng = np.random.default_rng(4)
    da, db = random_dense(rng, 6, 7, 0.3), random_dense(rng, 5, 4, 0.4)
    ar, ac = np.nonzero(da)
    br, bc = np.nonzero(db)
    r, c = oracle.kron(ar, ac, br, bc, 5, 4)
    er, ec = np.nonzero(np.kron(da, db))
    assert np.array_equal(r, er) and np.array_equal(c, ec)


def test_transpose_matches_dense():
    rng = np.random.default_rng(12)
    d = random_dense(rng, 70, 45, 0.05)
    A = dcsr_from_dense(d)
    T = oracle.transpose(A)
    assert (T.nrows, T.ncols) == (45, 70)
    assert np.array_equal(dense_from_dcsr(T, 45, 70), d.T)
    back = oracle.transpose(T)
    assert np.array_equal(back.rpt, A.rpt) and np.array_equal(back.rows, A.rows) and np.array_equal(back.cols, A.cols)
    # a planted mistake (no transpose) fails the dense check
    assert not np.array_equal(dense_from_dcsr(A, 70, 70)[:45, :70], d.T)
    assert oracle.transpose(dcsr_from_dense(np.zeros((3, 4), np.uint8))).nnz == 0


@pytest.mark.parametrize("i,j,nrows,ncols", [(0, 0, 60, 50), (10, 5, 20, 30), (59, 49, 1, 1), (0, 0, 0, 0),
                                             (30, 0, 0, 10), (5, 7, 100, 100), (60, 0, 5, 5)])
def test_submatrix_matches_dense_slice(i, j, nrows, ncols):
    rng = np.random.default_rng(13)
    d = random_dense(rng, 60, 50, 0.08)
    S = oracle.submatrix(dcsr_from_dense(d), i, j, nrows, ncols)
    want = d[i:i + nrows, j:j + ncols]
    got = dense_from_dcsr(S, want.shape[0], want.shape[1]) if S.nnz else np.zeros_like(want)
    assert np.array_equal(got, want)
    # canonical: no empty rows, rows ascending
    assert np.all(np.diff(S.rpt.astype(np.int64)) > 0) and np.all(np.diff(S.rows.astype(np.int64)) > 0)


def test_reduce_is_row_any():
    rng = np.random.default_rng(14)
    d = random_dense(rng, 80, 30, 0.02)
    R = oracle.reduce(dcsr_from_dense(d))
    assert np.array_equal(dense_from_dcsr(R, 80, 1)[:, 0], d.any(axis=1).astype(np.uint8))
    assert np.all(R.cols == 0)


def test_exclusive_scan_closed_form():
    rng = np.random.default_rng(6)
    a = rng.integers(0, 4, 100001).astype(np.uint32)
    s, t = oracle.exclusive_scan(a)
    assert t == int(a.sum())
    assert np.array_equal(s.astype(np.int64), np.r_[0, np.cumsum(a.astype(np.int64))[:-1]])


def test_csr_from_pairs_known_answer():
    ka = _load("known_answers.json")["csr_cpu_from_pairs"]
    p = np.array(ka["pairs"], np.uint32)
    rpt, cols = oracle.csr_from_pairs(p[:, 0], p[:, 1], ka["m"])
    assert rpt.tolist() == ka["rpt"] and cols.tolist() == ka["cols"]


def test_dcsr_csr_round_trip():
    rng = np.random.default_rng(7)
    d = random_dense(rng, 120, 50, 0.03)
    A = dcsr_from_dense(d)
    csr_rpt = oracle.dcsr_to_csr_rpt(A, 120)
    assert np.array_equal(csr_rpt.astype(np.int64), np.r_[0, np.cumsum(d.sum(axis=1))])
    rpt, rows = oracle.csr_to_dcsr(csr_rpt, 120)
    assert np.array_equal(rpt, A.rpt) and np.array_equal(rows, A.rows)


def test_coo_build_known_answer_and_sort():
    ka = _load("known_answers.json")["coo_constructor"]
    r, c = oracle.coo_build(ka["rows"], ka["cols"], sorted_=True, no_duplicates=False)
    assert r.size == ka["nnz_after"]
    assert [[int(x), int(y)] for x, y in zip(r, c)] == ka["pairs_after"]
    rng = np.random.default_rng(8)
    rr = rng.integers(0, 1000, 5000).astype(np.uint32)
    cc = rng.integers(0, 1000, 5000).astype(np.uint32)
    r, c = oracle.coo_build(rr, cc, sorted_=False, no_duplicates=True)
    order = np.lexsort((cc, rr))
    assert np.array_equal(r, rr[order]) and np.array_equal(c, cc[order])


# ---- MLP ----------------------------------------------------------------------
def test_mlp_gradients_finite_differences():
    F, H1, H2 = 5, 8, 6
    rng = np.random.default_rng(9)
    p = (rng.standard_normal(oracle.mlp_param_count(F, H1, H2)) * 0.5).astype(np.float32)
    x = rng.standard_normal((13, F)).astype(np.float32)
    t = rng.standard_normal(13).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(p, x, t, F, H1, H2)
    y = oracle.mlp_forward(p, x, F, H1, H2)
    assert abs(loss - float(np.mean(0.5 * (y.astype(np.float64) - t) ** 2))) < 1e-5
    eps = 1e-2
    for i in range(0, p.size, 3):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = oracle.mlp_loss_grad(pp, x, t, F, H1, H2)
        lm, _ = oracle.mlp_loss_grad(pm, x, t, F, H1, H2)
        num = (lp - lm) / (2 * eps)
        assert abs(num - g[i]) < 2e-3 + 2e-2 * abs(num), i


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    m, v = np.zeros(3, np.float32), np.zeros(3, np.float32)
    g = np.array([0.3, -0.1, 0.0], np.float32)
    oracle.adam_step(p, m, v, g, 0.01, 0.9, 0.999, 1e-8, 1)
    # first step: m_hat = g, v_hat = g^2, update = lr * g / (|g| + eps)
    expect = np.array([1.0 - 0.01, -2.0 + 0.01, 0.5], np.float32)
    assert np.allclose(p, expect, atol=1e-6)


# ---- row digests ----------------------------------------------------------------
def _splitmix64_np(x):
    """splitmix64 finaliser (Vigna), restated in numpy uint64 arithmetic."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _row_hashes_np(m: oracle.Dcsr):
    mixed = _splitmix64_np(m.cols)
    rpt = m.rpt.astype(np.int64)
    out = np.zeros(m.nzr, np.uint64)
    for q in range(m.nzr):
        with np.errstate(over="ignore"):
            out[q] = np.sum(mixed[rpt[q]:rpt[q + 1]], dtype=np.uint64)
    return out


def test_splitmix64_known_values():
    # splitmix64 as a generator from state 0: its first outputs are the
    # finaliser applied to k * golden-gamma (published reference values)
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        got = _splitmix64_np(np.array([0, 1], np.uint64) * g)
    assert [int(x) for x in got] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]


def test_spgemm_digest_matches_product_rows():
    rng = np.random.default_rng(21)
    a = random_dense(rng, 90, 70, 0.06)
    b = random_dense(rng, 70, 3000, 0.02)
    A, B = dcsr_from_dense(a), dcsr_from_dense(b)
    ref = dcsr_from_dense(oracle.dense_bool_mm(a, b))
    cnt, h = oracle.spgemm_digest(A, B)
    # stored rows of A with an empty product get (0, 0); the others match the
    # brute-force product's rows in order
    nz = cnt > 0
    assert np.array_equal(A.rows[nz], ref.rows)
    assert np.array_equal(cnt[nz].astype(np.int64), np.diff(ref.rpt.astype(np.int64)))
    assert np.array_equal(h[nz], _row_hashes_np(ref))
    assert np.all(h[~nz] == 0)
    # a row range of A gives the matching slice
    c2, h2 = oracle.spgemm_digest(A, B, 10, 40)
    assert np.array_equal(c2, cnt[10:40]) and np.array_equal(h2, h[10:40])
    # the GPU-side digest of the canonical product agrees, also from chunks
    hd, bad = oracle.dcsr_row_digest(ref.rpt, ref.cols)
    assert bad == 0 and np.array_equal(hd, h[nz])
    k = int(ref.rpt[20])
    hc, bad = oracle.dcsr_row_digest(ref.rpt, ref.cols[k:], base=k, r0=20, r1=ref.nzr)
    assert bad == 0 and np.array_equal(hc, hd[20:])


def test_digest_planted_mistakes_are_caught():
    rng = np.random.default_rng(22)
    a = random_dense(rng, 60, 60, 0.08)
    A = dcsr_from_dense(a)
    ref = dcsr_from_dense(oracle.dense_bool_mm(a, a))
    h, _ = oracle.dcsr_row_digest(ref.rpt, ref.cols)
    # a changed column (same count) changes that row's digest
    cols = ref.cols.copy()
    cols[ref.rpt[3]] ^= 1
    h2, _ = oracle.dcsr_row_digest(ref.rpt, cols)
    assert h2[3] != h[3] and np.array_equal(np.delete(h2, 3), np.delete(h, 3))
    # two swapped columns keep the digest but are reported as unsorted
    r = int(np.flatnonzero(np.diff(ref.rpt.astype(np.int64)) >= 2)[0])
    cols = ref.cols.copy()
    s = int(ref.rpt[r])
    cols[s], cols[s + 1] = cols[s + 1], cols[s]
    h3, bad = oracle.dcsr_row_digest(ref.rpt, cols)
    assert bad == 1 and np.array_equal(h3, h)
    # a product with a transposed operand has different digests
    _, ht = oracle.spgemm_digest(A, dcsr_from_dense(np.ascontiguousarray(a.T)))
    _, hp = oracle.spgemm_digest(A, A)
    assert not np.array_equal(ht, hp)


# ---- MLP forward against an independent evaluation -------------------------------
def _mlp_numpy(p, x, F, H1, H2, relu1=True, relu2=True):
    """relu(relu(x W1^T + b1) W2^T + b2) . w3 + b3, parameters laid out
    W1[H1][F], b1[H1], W2[H2][H1], b2[H2], w3[H2], b3 (the layout that
    include/clbool_c.h documents for CLB_MLP_PARAMS)."""
    p = p.astype(np.float64)
    o = np.cumsum([0, H1 * F, H1, H2 * H1, H2, H2])
    W1, b1 = p[o[0]:o[1]].reshape(H1, F), p[o[1]:o[2]]
    W2, b2 = p[o[2]:o[3]].reshape(H2, H1), p[o[3]:o[4]]
    w3, b3 = p[o[4]:o[5]], p[o[5]]
    h = x.astype(np.float64) @ W1.T + b1
    h = np.maximum(h, 0) if relu1 else h
    h = h @ W2.T + b2
    h = np.maximum(h, 0) if relu2 else h
    return h @ w3 + b3


@pytest.mark.parametrize("F,H1,H2", [(5, 32, 32), (5, 8, 6), (3, 4, 7)])
def test_mlp_forward_matches_numpy(F, H1, H2):
    rng = np.random.default_rng(31 + H1)
    n = oracle.mlp_param_count(F, H1, H2)
    assert n == H1 * F + H1 + H2 * H1 + H2 + H2 + 1
    p = (rng.standard_normal(n) * 0.7).astype(np.float32)
    x = rng.standard_normal((200, F)).astype(np.float32)
    y = oracle.mlp_forward(p, x, F, H1, H2).astype(np.float64)
    ref = _mlp_numpy(p, x, F, H1, H2)
    assert np.allclose(y, ref, rtol=1e-5, atol=1e-5)
    # planted mistakes: either ReLU dropped, or W1 read transposed, must differ
    for wrong in (_mlp_numpy(p, x, F, H1, H2, relu1=False), _mlp_numpy(p, x, F, H1, H2, relu2=False)):
        assert np.max(np.abs(y - wrong)) > 1e-2
    if F != H1:
        q = p.copy()
        q[:H1 * F] = p[:H1 * F].reshape(F, H1).T.reshape(-1)  # W1 read as [F][H1]
        assert np.max(np.abs(y - _mlp_numpy(q, x, F, H1, H2))) > 1e-2
    # the production parameter count
    if (F, H1, H2) == (5, 32, 32):
        assert n == 1281
