"""Parity of the HIP boolean SpGEMM with the oracle (through the C-ABI).

Mirrors test/test_operations.cpp:127-134 (multiplication_hash) and the
configurations of the benchmark workloads at sizes the oracle finishes in
seconds; every output is compared bit for bit (rpt, rows, cols).
"""
import json
import os

import numpy as np
import pytest

import clbool_amd as cb
from clbool_amd import autotune as at
import oracle
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr, upload

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_readme_example(controls):
    ka = json.load(open(os.path.join(GOLD, "known_answers.json")))["readme_product"]
    _, a = pairs_dcsr(controls, ka["a"]["rows"], ka["a"]["cols"], 5, 5)
    _, b = pairs_dcsr(controls, ka["b"]["rows"], ka["b"]["cols"], 5, 5)
    c = cb.dcsr.matrix_multiplication_hash(controls, a, b)
    rpt, rows, cols = c.to_host(controls)
    assert rpt.tolist() == ka["c"]["rpt"] and rows.tolist() == ka["c"]["rows"] and cols.tolist() == ka["c"]["cols"]
    assert (c.nrows, c.ncols) == (5, 5)


@pytest.mark.parametrize("k", [50, 55])
@pytest.mark.parametrize("size", [3000, 3200, 3400, 3600, 3800])
def test_multiplication_hash_reference_sizes(controls, size, k):
    # test_multiplication_hash (test/src/operations.cpp:43-81): A = A * A
    r, c = oracle.generate_coo_pairs(size * k, size)
    A, a = pairs_dcsr(controls, r, c, size, size)
    got = cb.dcsr.matrix_multiplication_hash(controls, a, a)
    assert_dcsr_equal(controls, got, oracle.spgemm_reference(A, A), f"{size}x{k}")


@pytest.mark.parametrize("seed", range(1, 9))
def test_uniform_1024_10k_pairs(controls, seed):
    ra, ca = gen.uniform_pairs(10000, 1024, 1024, seed)
    rb, cb_ = gen.uniform_pairs(10000, 1024, 1024, seed + 1)
    A, a = pairs_dcsr(controls, ra, ca, 1024, 1024)
    B, b = pairs_dcsr(controls, rb, cb_, 1024, 1024)
    got, st = cb.dcsr.spgemm(controls, a, b)
    ref = oracle.spgemm_reference(A, B)
    assert_dcsr_equal(controls, got, ref, f"seed {seed}")
    assert st["products"] == int(oracle.workload(A, B).sum())
    assert st["nnz"] == ref.nnz and st["nzr"] == ref.nzr


@pytest.mark.parametrize("scale", [10, 12, 14, 16])
def test_rmat_power_law(controls, scale):
    r, c = gen.rmat(scale, 16, scale)
    n = 1 << scale
    A, a = pairs_dcsr(controls, r, c, n, n)
    got, st = cb.dcsr.spgemm(controls, a, a)
    assert_dcsr_equal(controls, got, oracle.spgemm(A, A), f"rmat{scale}")
    assert st["products"] == int(oracle.workload(A, A).sum())


def test_rectangular_and_wide_columns(controls):
    # A: 3000 x 2^22, B: 2^22 x 2^23 -- bitmap windows are sliced
    ra, ca = gen.uniform_pairs(60000, 3000, 1 << 22, 11)
    hub = np.repeat(np.arange(0, 3000, 3, dtype=np.uint32), 30)
    ra2 = np.r_[ra, hub]
    ca2 = np.r_[ca, np.tile(np.arange(30, dtype=np.uint32), 1000)]
    key = np.unique((ra2.astype(np.uint64) << np.uint64(32)) | ca2)
    ra, ca = (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    rb = np.repeat(np.arange(30, dtype=np.uint32), 3000)
    cbv = np.random.default_rng(3).integers(0, 1 << 23, rb.size).astype(np.uint32)
    key = np.unique((rb.astype(np.uint64) << np.uint64(32)) | cbv)
    rb, cbv = (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    A, a = pairs_dcsr(controls, ra, ca, 3000, 1 << 22)
    B, b = pairs_dcsr(controls, rb, cbv, 1 << 22, 1 << 23)
    got = cb.dcsr.matrix_multiplication_hash(controls, a, b)
    assert (got.nrows, got.ncols) == (3000, 1 << 23)
    assert_dcsr_equal(controls, got, oracle.spgemm(A, B), "wide")


def test_hypersparse_2p24(controls):
    r, c = gen.hypersparse_zipf(seed=1, log_n=24, row_frac=0.01)
    n = 1 << 24
    A, a = pairs_dcsr(controls, r, c, n, n)
    got, st = cb.dcsr.spgemm(controls, a, a)
    assert_dcsr_equal(controls, got, oracle.spgemm(A, A), "hypersparse")


@pytest.mark.parametrize("knobs", [
    dict(hash_max_slots=256, bitmap_min_products=1 << 30, bitmap_slot_ratio=1, tiny_max_products=0),
    dict(hash_max_slots=16384, bitmap_min_products=1 << 30, bitmap_slot_ratio=1, tiny_max_products=64),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0),
    dict(hash_max_slots=1024, bitmap_min_products=4096, bitmap_slot_ratio=4, tiny_max_products=16),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0, bitmap_persistent=1),
    dict(hash_max_slots=1024, bitmap_min_products=4096, bitmap_slot_ratio=4, tiny_max_products=16,
         bitmap_persistent=1),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0,
         gather_combine=1),
    dict(hash_max_slots=1024, bitmap_min_products=4096, bitmap_slot_ratio=4, tiny_max_products=16,
         gather_combine=1),
    dict(hash_max_slots=16384, bitmap_min_products=1 << 30, bitmap_slot_ratio=1, tiny_max_products=0,
         **{f"hash_wide_{b}": 1 for b in range(7, 14)}),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0,
         bitmap_window_bits=13),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0,
         bitmap_window_bits=15),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0,
         emit_sparse_bits=1),
    dict(hash_max_slots=16384, bitmap_min_products=1, bitmap_slot_ratio=1, tiny_max_products=0,
         emit_sparse_bits=1 << 20, dump_bits_per_chunk=65536),
    # the ratio rule's bitmap rows of the wide case go to the largest tables
    dict(hash_max_slots=16384, bitmap_min_products=1 << 30, bitmap_slot_ratio=4, tiny_max_products=0,
         tables_over_slices=1),
])
def test_every_kernel_path(controls, knobs):
    """Force the rows through each bin family (hash only, tiny, bitmap only)."""
    old = controls.tuning()
    t = controls.tuning()
    for k, v in knobs.items():
        at._set(t, k, v)
    controls.set_tuning(t)
    try:
        for scale, ncols in ((12, 1 << 12), (11, 1 << 22)):
            n = 1 << scale
            A, a = pairs_dcsr(controls, *gen.rmat(scale, 8, 100 + scale), n, n)
            rb, cbv = gen.rmat(scale, 8, 200 + scale)
            if ncols != n:  # spread B's columns over a wide range (sliced bitmaps)
                cbv = (cbv.astype(np.uint64) * 4099 % ncols).astype(np.uint32)
                key = np.unique((rb.astype(np.uint64) << np.uint64(32)) | cbv)
                rb, cbv = (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            B, b = pairs_dcsr(controls, rb, cbv, n, ncols)
            got = cb.dcsr.matrix_multiplication_hash(controls, a, b)
            assert_dcsr_equal(controls, got, oracle.spgemm(A, B), f"{knobs} ncols={ncols}")
    finally:
        controls.set_tuning(old)


def test_empty_operands_and_shapes(controls):
    r, c = gen.uniform_pairs(500, 100, 80, 3)
    A, a = pairs_dcsr(controls, r, c, 100, 80)
    z = cb.matrix_dcsr(80, 60)
    out = cb.dcsr.matrix_multiplication_hash(controls, a, z)
    assert out.nnz == 0 and (out.nrows, out.ncols) == (100, 60)   # hash.cpp:64-67
    out = cb.dcsr.matrix_multiplication_hash(controls, cb.matrix_dcsr(7, 100), a)
    assert out.nnz == 0 and (out.nrows, out.ncols) == (7, 80)
    # no products at all: A's columns miss every B row
    Bm, b = pairs_dcsr(controls, [79], [3], 80, 60)
    Am, a2 = pairs_dcsr(controls, [0, 5], [1, 2], 100, 80)
    out = cb.dcsr.matrix_multiplication_hash(controls, a2, b)
    assert out.nnz == 0 and (out.nrows, out.ncols) == (100, 60)


def test_dimension_mismatch_raises(controls):
    _, a = pairs_dcsr(controls, [0], [0], 4, 5)
    _, b = pairs_dcsr(controls, [0], [0], 4, 5)
    with pytest.raises(cb.ClboolError) as e:
        cb.dcsr.matrix_multiplication_hash(controls, a, b)     # hash.cpp:56-61
    assert e.value.status_name == "CLBOOL_INVALID_ARGUMENT"


def test_identity(controls):
    r, c = gen.rmat(12, 8, 77)
    n = 1 << 12
    A, a = pairs_dcsr(controls, r, c, n, n)
    I, i = pairs_dcsr(controls, np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), n, n)
    assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication_hash(controls, i, a), A, "I*A")
    assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication_hash(controls, a, i), A, "A*I")


def test_row_blocks_concatenate_to_full(controls):
    r, c = gen.rmat(13, 16, 5)
    n = 1 << 13
    A, a = pairs_dcsr(controls, r, c, n, n)
    full = oracle.spgemm(A, A)
    bounds = [0, 17, A.nzr // 3, A.nzr // 2, A.nzr]
    parts = []
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        blk, st = cb.dcsr.spgemm(controls, a, a, row_begin=lo, row_end=hi)
        parts.append(blk.to_host(controls) + (blk.nnz,))
    cols = np.concatenate([p[2][:p[3]] for p in parts])
    rows = np.concatenate([p[1] for p in parts])
    assert np.array_equal(cols, full.cols) and np.array_equal(rows, full.rows)


def test_repeated_calls_reuse_memory(controls):
    r, c = gen.rmat(12, 16, 9)
    n = 1 << 12
    A, a = pairs_dcsr(controls, r, c, n, n)
    ref = oracle.spgemm(A, A)
    for _ in range(5):
        assert_dcsr_equal(controls, cb.dcsr.matrix_multiplication_hash(controls, a, a), ref, "repeat")


def _hub_product(seed, nrows=1200, hubs=64, ncols=1 << 20):
    """A picks 1..30 of `hubs` B rows whose lengths range 300..40000 over
    2^20 columns: bitmap rows from sparse to dense in one product."""
    rng = np.random.default_rng(seed)
    lens = rng.choice([300, 3000, 12000, 40000], hubs)
    rb = np.repeat(np.arange(hubs, dtype=np.uint32), lens)
    cbv = rng.integers(0, ncols, rb.size).astype(np.uint32)
    key = np.unique((rb.astype(np.uint64) << np.uint64(32)) | cbv)
    rb, cbv = (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    k = rng.integers(1, 31, nrows)
    ra = np.repeat(np.arange(nrows, dtype=np.uint32), k)
    ca = rng.integers(0, hubs, ra.size).astype(np.uint32)
    key = np.unique((ra.astype(np.uint64) << np.uint64(32)) | ca)
    ra, ca = (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return (ra, ca, nrows, hubs), (rb, cbv, hubs, ncols)


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("dense_knob", [16, 256, 65536])
def test_dense_bitmap_rows(controls, dense_knob, persistent):
    """Rows over 2^20 columns (one symbolic bitmap window, two numeric ones):
    dumped dense bitmaps (expanded by the emission kernel) and in-LDS emission
    must give the same exact result."""
    (ra, ca, m, k), (rb, cbv, k2, n) = _hub_product(5)
    A, a = pairs_dcsr(controls, ra, ca, m, k)
    B, b = pairs_dcsr(controls, rb, cbv, k2, n)
    ref = oracle.spgemm(A, B)
    old = controls.tuning()
    t = controls.tuning()
    t.dump_bits_per_chunk = dense_knob
    t.bitmap_persistent = persistent
    controls.set_tuning(t)
    try:
        got, st = cb.dcsr.spgemm(controls, a, b, timing=True)
        assert_dcsr_equal(controls, got, ref, f"plain dense_knob={dense_knob} persistent={persistent}")
        emits = st["launches"][4 * 1 + 3]
        assert emits == (0 if dense_knob == 65536 else 1)
        cb.predictor_reset(controls, 3)
        got, st = cb.dcsr.spgemm(controls, a, b, mode=cb.SPGEMM_LEARNED, timing=True)
        assert_dcsr_equal(controls, got, ref, f"learned dense_knob={dense_knob} persistent={persistent}")
    finally:
        controls.set_tuning(old)


@pytest.mark.parametrize("ncols", [1 << 22, (1 << 20) + 12345])
@pytest.mark.parametrize("dense_knob", [4, 64, 65536])
def test_dense_rows_over_column_slices(controls, dense_knob, ncols):
    """Columns over several bitmap windows of 2^20 columns (2^22: four;
    2^20 + 12345: two, the last one partial): rows whose first slice is dense
    are dumped slice by slice and expanded by the emission kernel; the others
    are emitted slice by slice in LDS.  Both passes and the learned pass must
    agree with the oracle."""
    (ra, ca, m, k), (rb, cbv, k2, n) = _hub_product(6, nrows=600, ncols=ncols)
    A, a = pairs_dcsr(controls, ra, ca, m, k)
    B, b = pairs_dcsr(controls, rb, cbv, k2, n)
    ref = oracle.spgemm(A, B)
    old = controls.tuning()
    t = controls.tuning()
    t.dump_bits_per_chunk = dense_knob
    controls.set_tuning(t)
    try:
        got, st = cb.dcsr.spgemm(controls, a, b, timing=True)
        assert_dcsr_equal(controls, got, ref, f"plain sliced dense_knob={dense_knob}")
        assert (st["dense_rows"] > 0) == (dense_knob != 65536)
        cb.predictor_reset(controls, 4)
        got, st = cb.dcsr.spgemm(controls, a, b, mode=cb.SPGEMM_LEARNED, timing=True)
        assert_dcsr_equal(controls, got, ref, f"learned sliced dense_knob={dense_knob}")
    finally:
        controls.set_tuning(old)
