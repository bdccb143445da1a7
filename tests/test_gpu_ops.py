"""Parity of the HIP additions, Kronecker product, conversions, scan and
pair sort with the oracle (mirrors test/test_operations.cpp,
test/test_procedures.cpp, test/test_conversions.cpp, test/test_errors.cpp)."""
import json
import os

import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- COO addition: addition_coo_{small,medium,large} (test_operations.cpp:142-176)
@pytest.mark.parametrize("size,k_a,k_b", [(0, 3, 3), (15, 0, 4), (15, 4, 0), (30, 9, 9), (1000, 10, 25),
                                          (3800, 25, 10), (4000, 30, 65), (14500, 65, 30)])
def test_addition_coo(controls, size, k_a, k_b):
    ar, ac = oracle.generate_coo_pairs(size * k_a, size) if size else (np.zeros(0, np.uint32),) * 2
    br, bc = oracle.generate_coo_pairs(size * k_b, size) if size else (np.zeros(0, np.uint32),) * 2
    a = cb.matrix_coo.from_host(controls, ar, ac, size, size)
    b = cb.matrix_coo.from_host(controls, br, bc, size, size)
    out = cb.coo.matrix_addition(controls, a, b)
    er, ec = oracle.coo_add(ar, ac, br, bc)
    assert out.nnz == er.size
    r, c = out.to_host(controls)
    assert np.array_equal(r, er) and np.array_equal(c, ec)


def test_addition_coo_shape_mismatch(controls):
    a = cb.matrix_coo.from_host(controls, [0], [0], 3, 3)
    b = cb.matrix_coo.from_host(controls, [0], [0], 3, 4)
    with pytest.raises(cb.ClboolError) as e:
        cb.coo.matrix_addition(controls, a, b)
    assert e.value.status_name == "CLBOOL_INVALID_ARGUMENT"


# ---- CSR addition (test_operations.cpp:182-217, operations.cpp:420-452)
@pytest.mark.parametrize("size,k_a,k_b", [(15, 0, 3), (15, 3, 0), (300, 9, 1), (1000, 10, 25), (7000, 20, 35)])
def test_addition_csr(controls, size, k_a, k_b):
    ar, ac = oracle.generate_coo_pairs(size * k_a, size)
    br, bc = oracle.generate_coo_pairs(size * k_b, size)
    arpt, acols = oracle.csr_from_pairs(ar, ac, size)
    brpt, bcols = oracle.csr_from_pairs(br, bc, size)
    a = cb.csr_from_cpu(controls, arpt, acols, size, size)
    b = cb.csr_from_cpu(controls, brpt, bcols, size, size)
    out = cb.csr.matrix_addition(controls, a, b)
    er, ec = oracle.coo_add(ar, ac, br, bc)
    erpt, ecols = oracle.csr_from_pairs(er, ec, size)
    assert out.nnz == ecols.size
    rpt, cols = out.to_host(controls)
    assert np.array_equal(rpt, erpt) and np.array_equal(cols, ecols)


# ---- transpose / submatrix / reduce (test_operations.cpp:10-66,
# operations.cpp:84-238)
def _dcsr_of(controls, r, c, m, n):
    A = oracle.pairs_to_dcsr(r, c, m, n)
    return A, cb.matrix_dcsr_from_cpu(controls, A.rpt, A.rows, A.cols, m, n)


@pytest.mark.parametrize("size,k", [(10, 1), (100, 5), (1000, 10), (5000, 50), (10000, 30)])
def test_transpose(controls, size, k):
    r, c = oracle.generate_coo(size * k, size)  # generate_coo_cpu draws in [0, size]
    A, a = _dcsr_of(controls, r, c, size + 1, size + 1)
    got = cb.dcsr.transpose(controls, a)
    assert (got.nrows, got.ncols) == (size + 1, size + 1)
    assert_dcsr_equal(controls, got, oracle.transpose(A), f"transpose size={size} k={k}")


def test_transpose_rectangular_and_empty(controls):
    r, c = gen.uniform_pairs(20000, 300, 7000, 5)
    A, a = _dcsr_of(controls, r, c, 300, 7000)
    got = cb.dcsr.transpose(controls, a)
    assert (got.nrows, got.ncols) == (7000, 300)
    assert_dcsr_equal(controls, got, oracle.transpose(A), "transpose 300x7000")
    e = cb.dcsr.transpose(controls, cb.matrix_dcsr(5, 9))
    assert (e.nrows, e.ncols, e.nnz) == (9, 5, 0)


@pytest.mark.parametrize("size,k", [(100, 10), (1000, 10), (3000, 50), (10000, 5)])
def test_submatrix(controls, size, k):
    ar, ac = oracle.generate_coo_pairs(size * k, size)
    A, a = _dcsr_of(controls, ar, ac, size, size)
    rng = np.random.default_rng(size + k)
    for trial in range(4):
        i1, i2, i3, i4 = (int(x) for x in rng.integers(0, size + 1, 4))
        i, nrows, j, ncols = min(i1, i2), abs(i1 - i2), min(i3, i4), abs(i3 - i4)
        want = oracle.submatrix(A, i, j, nrows, ncols)
        got = cb.dcsr.submatrix(controls, a, i, j, nrows, ncols)
        if want.nnz == 0:
            assert (got.nnz, got.nrows, got.ncols) == (0, 0, 0)
        else:
            assert (got.nrows, got.ncols) == (nrows, ncols)
            assert_dcsr_equal(controls, got, want, f"submatrix {i},{j},{nrows},{ncols}")


def test_submatrix_zero_and_edges(controls):
    ar, ac = oracle.generate_coo_pairs(5000, 500)
    A, a = _dcsr_of(controls, ar, ac, 500, 500)
    for i, j, nrows, ncols in [(0, 0, 0, 0), (0, 0, 500, 500), (499, 499, 1, 1), (0, 250, 500, 0),
                               (100, 100, 2 ** 32 - 101, 2 ** 32 - 101)]:
        want = oracle.submatrix(A, i, j, nrows, ncols)
        got = cb.dcsr.submatrix(controls, a, i, j, nrows, ncols)
        if want.nnz == 0:
            assert (got.nnz, got.nrows, got.ncols) == (0, 0, 0)
        else:
            assert_dcsr_equal(controls, got, want, f"submatrix {i},{j},{nrows},{ncols}")


@pytest.mark.parametrize("size,k", [(100, 1), (1000, 10), (30000, 3)])
def test_reduce(controls, size, k):
    r, c = oracle.generate_coo(size * k, size)
    A, a = _dcsr_of(controls, r, c, size + 1, size + 1)
    got = cb.dcsr.reduce(controls, a)
    assert (got.nrows, got.ncols) == (size + 1, 1)
    assert_dcsr_equal(controls, got, oracle.reduce(A), f"reduce size={size}")
    assert cb.dcsr.reduce(controls, cb.matrix_dcsr(4, 4)).nnz == 0


# ---- DCSR union (the closure's X += T)
def test_addition_dcsr(controls):
    r1, c1 = gen.rmat(12, 8, 1)
    r2, c2 = gen.rmat(12, 8, 2)
    n = 1 << 12
    A, a = pairs_dcsr(controls, r1, c1, n, n)
    B, b = pairs_dcsr(controls, r2, c2, n, n)
    out = cb.dcsr.addition(controls, a, b)
    er, ec = oracle.coo_add(r1, c1, r2, c2)
    assert_dcsr_equal(controls, out, oracle.pairs_to_dcsr(er, ec), "dcsr union")
    empty = cb.matrix_dcsr(n, n)
    assert_dcsr_equal(controls, cb.dcsr.addition(controls, a, empty), A, "A + 0")


# ---- Kronecker: kronecker_dcsr_* (test_operations.cpp:251-273)
@pytest.mark.parametrize("k", [0, 1, 2, 4, 6])
def test_kronecker_dcsr(controls, k):
    nnz = 1000 * k
    size_a, size_b = 10000, 10000
    ar, ac = oracle.generate_coo_pairs(nnz, size_a)
    A, a = pairs_dcsr(controls, ar, ac, size_a, size_a) if nnz else (None, cb.matrix_dcsr(size_a, size_a))
    if k >= 4:
        # keep the product under the uint32 nonzero limit of the reference's matrices
        br, bc = oracle.generate_coo_pairs(500 * k, 400)
        B, b = pairs_dcsr(controls, br, bc, 400, 400)
        size_b = 400
    else:
        br, bc = ar, ac
        b = a
    out = cb.dcsr.kronecker_product(controls, a, b)
    assert (out.nrows, out.ncols) == ((size_a * size_b) & 0xFFFFFFFF, (size_a * size_b) & 0xFFFFFFFF) or nnz == 0
    er, ec = oracle.kron(ar, ac, br, bc, size_b, size_b)
    assert_dcsr_equal(controls, out, oracle.pairs_to_dcsr(er, ec), f"kron k={k}")


def test_kronecker_automaton_times_graph(controls):
    # 64 x 64 automaton (200 uniform nnz) (x) an R-MAT graph, scaled down
    mr, mc = gen.uniform_pairs(200, 64, 64, 4)
    gr, gc = gen.rmat(12, 8, 4)
    M, m = pairs_dcsr(controls, mr, mc, 64, 64)
    G, g = pairs_dcsr(controls, gr, gc, 1 << 12, 1 << 12)
    out = cb.dcsr.kronecker_product(controls, m, g)
    er, ec = oracle.kron(mr, mc, gr, gc, 1 << 12, 1 << 12)
    assert_dcsr_equal(controls, out, oracle.pairs_to_dcsr(er, ec), "automaton kron")


def test_kronecker_many_rows(controls):
    """More output tiles than one launch takes (5M A rows times a B of one
    nonzero and of three rows): the fill kernel runs in several launches of
    whole A rows."""
    n = 5 << 20
    ar = np.arange(n, dtype=np.uint32)
    ac = (ar * 7) % n
    for br, bc in ((np.array([3], np.uint32), np.array([2], np.uint32)),
                   (np.array([0, 0, 5], np.uint32), np.array([1, 4, 0], np.uint32))):
        A, a = pairs_dcsr(controls, ar, ac, n, n)
        B, b = pairs_dcsr(controls, br, bc, 8, 8)
        out = cb.dcsr.kronecker_product(controls, a, b)
        er, ec = oracle.kron(ar, ac, br, bc, 8, 8)
        assert_dcsr_equal(controls, out, oracle.pairs_to_dcsr(er, ec), f"kron many rows, b nnz {br.size}")


# ---- conversions (test_conversions.cpp, conversions.cpp:49-84)
@pytest.mark.parametrize("size,k", [(200, 0), (200, 5), (4000, 30), (4700, 70), (29700, 40)])
def test_dcsr_csr_round_trips(controls, size, k):
    r, c = oracle.generate_coo_pairs(size * k, size)
    D = oracle.pairs_to_dcsr(r, c, size, size)
    d = cb.matrix_dcsr_from_cpu(controls, D.rpt, D.rows, D.cols, size)
    rpt, cols = oracle.csr_from_pairs(r, c, size)
    s = cb.csr_from_cpu(controls, rpt, cols, size, size)
    conv = cb.dcsr_to_csr(controls, d)
    if r.size:
        got_rpt, got_cols = conv.to_host(controls)
        assert np.array_equal(got_rpt, rpt) and np.array_equal(got_cols, cols)
    back = cb.csr_to_dcsr(controls, s)
    assert_dcsr_equal(controls, back, D, "csr->dcsr")


def test_coo_dcsr_conversions(controls):
    r, c = gen.rmat(13, 8, 8)
    n = 1 << 13
    coo_m = cb.matrix_coo.from_host(controls, r, c, n, n)
    D = oracle.pairs_to_dcsr(r, c)
    d = cb.coo_to_dcsr_shallow(controls, coo_m)
    assert_dcsr_equal(controls, d, D, "coo->dcsr")
    for conv in (cb.dcsr_to_coo_shallow, cb.dcsr_to_coo_deep):
        back = conv(controls, d)
        br, bc = back.to_host(controls)
        assert np.array_equal(br, r) and np.array_equal(bc, c)


# ---- procedures (test_procedures.cpp)
def test_coo_constructor_dedup(controls):
    ka = json.load(open(os.path.join(GOLD, "known_answers.json")))["coo_constructor"]
    m = cb.matrix_coo.from_host(controls, ka["rows"], ka["cols"], 7, 7, sorted=True, noDuplicates=False)
    assert m.nnz == ka["nnz_after"]
    r, c = m.to_host(controls)
    assert [[int(x), int(y)] for x, y in zip(r, c)] == ka["pairs_after"]


@pytest.mark.parametrize("size", [1, 15, 1000, 4995, 100000, 2_000_000])
def test_bitonic_sort_pairs(controls, size):
    # test_bitonic_sort (procedures.cpp:51-87): fill_random_matrix + std::sort
    rows = oracle.fastrandom_stream(size, 0, 1023, 2 * size).astype(np.uint32)
    rr, cc = rows[0::2].copy(), rows[1::2].copy()
    rb = cb.Buffer.from_array(controls, rr)
    cbf = cb.Buffer.from_array(controls, cc)
    cb.coo.sort_arrays(controls, rb, cbf, size)
    er, ec = oracle.coo_build(rr, cc, sorted_=False, no_duplicates=True)
    assert np.array_equal(rb.read(controls, size), er) and np.array_equal(cbf.read(controls, size), ec)


def test_unsorted_coo_constructor(controls):
    rng = np.random.default_rng(3)
    rr = rng.integers(0, 1 << 20, 300000).astype(np.uint32)
    cc = rng.integers(0, 1 << 24, 300000).astype(np.uint32)
    m = cb.matrix_coo.from_host(controls, rr, cc, 1 << 20, 1 << 24, sorted=False, noDuplicates=False)
    er, ec = oracle.coo_build(rr, cc, sorted_=False, no_duplicates=False)
    r, c = m.to_host(controls)
    assert m.nnz == er.size and np.array_equal(r, er) and np.array_equal(c, ec)


@pytest.mark.parametrize("size", [1, 511, 4096, 4097, 10000, 590000, 1190000, 9_000_001])
def test_pref_sum(controls, size):
    # test_pref_sum (procedures.cpp:7-49): values in {0..3} plus a trailing 0
    rng = np.random.default_rng(size)
    vec = np.r_[rng.integers(0, 4, size), 0].astype(np.uint32)
    buf = cb.Buffer.from_array(controls, vec)
    total = cb.prefix_sum(controls, buf, size + 1)
    expect, etotal = oracle.exclusive_scan(vec)
    assert total == etotal
    assert np.array_equal(buf.read(controls, size + 1), expect)


def test_fill_with(controls):
    buf = cb.Buffer.create(controls, 4 * 1000)
    cb.fill_with(controls, buf, 1000, 7)
    assert np.all(buf.read(controls, 1000) == 7)


# ---- errors (test_errors.cpp)
def test_invalid_device_id():
    with pytest.raises(cb.ClboolError) as e:
        cb.create_controls(0, 64)
    assert e.value.status_name == "CLBOOL_INITIALIZATION_ERROR"


def test_zero_size_buffer(controls):
    with pytest.raises(cb.ClboolError) as e:
        cb.Buffer.create(controls, 0)
    assert e.value.status_name == "CLBOOL_CREATE_BUFFER_ERROR"
