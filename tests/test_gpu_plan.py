"""Products planned as row blocks (clb_spgemm_plan_*) and the bound-only one
pass, against the oracle.

* Every mode's plan covers A's stored rows with consecutive blocks, each
  block's result equals the oracle's product of those rows, and each result
  holds at most block_bound nonzeros.
* A learned plan cut by capacities that are far too small re-cuts the block
  that exceeds the bound by the exact sizes it found, and still gives the
  exact product.
* The bound-only one pass (capacities min(P, ncols), no predictor) and the
  learned one pass with bitmap rows sized by the prediction are exact.
"""
import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import autotune as at
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr

pytestmark = pytest.mark.gpu


def _over_model():
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    p[-1] = 1.0  # predicts 2^32 - 1: capacities clamp to min(P, ncols)
    return p


def _check_plan(controls, plan, A, full, bound, what):
    rows_seen = 0
    k = 0
    while k < plan.nblocks:
        lo, hi = plan.block(k)
        assert lo == rows_seen, f"{what}: block {k} starts at {lo}, expected {rows_seen}"
        C, st = plan.run(k)
        lo2, hi2 = plan.block(k)  # a re-cut block shrinks in place
        assert lo2 == lo
        ref = oracle.spgemm(A, A, row_sel=np.arange(lo, hi2, dtype=np.uint32))
        assert_dcsr_equal(controls, C, ref, f"{what} block {k} rows [{lo}, {hi2})")
        assert C.nnz <= bound or hi2 - lo == 1, f"{what}: block {k} holds {C.nnz} > {bound}"
        assert st["products"] == int(oracle.workload(
            oracle.Dcsr(*_rows(A, lo, hi2), A.nrows, A.ncols), A).sum())
        rows_seen = hi2
        k += 1
    assert rows_seen == A.nzr
    return k


def _rows(A, lo, hi):
    rpt = A.rpt[lo:hi + 1].astype(np.int64)
    return ((rpt - rpt[0]).astype(np.uint32), A.rows[lo:hi].copy(), A.cols[rpt[0]:rpt[-1]].copy())


@pytest.mark.parametrize("mode", [cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED, cb.SPGEMM_BOUND])
def test_plan_blocks_match_oracle(controls, mode):
    n = 1 << 13
    A, a = pairs_dcsr(controls, *gen.rmat(13, 16, 41), n, n)
    full = oracle.spgemm(A, A)
    cb.predictor_set_params(controls, _over_model())
    bound = max(full.nnz // 7, n)
    plan = cb.dcsr.plan(controls, a, a, mode=mode, block_bound=bound)
    info = plan.info()
    assert info["products"] == int(oracle.workload(A, A).sum())
    if mode == cb.SPGEMM_PLAIN:
        assert info["planned_nnz"] == full.nnz
    else:
        assert info["planned_nnz"] == int(np.minimum(oracle.workload(A, A), n).sum())
    nb = _check_plan(controls, plan, A, full, bound, f"mode {mode}")
    assert nb >= 7
    # the default bound (2^32 - 1) takes the whole product in one block
    one = cb.dcsr.plan(controls, a, a, mode=mode)
    assert one.nblocks == 1 and one.block(0) == (0, A.nzr)
    C, _ = one.run(0)
    assert_dcsr_equal(controls, C, full, f"mode {mode}, one block")


def test_learned_plan_recuts_blocks_over_the_bound(controls):
    # an all-zero model predicts 0 columns for every row: capacities are
    # CAP_PAD (8) for hash rows, so the cut by capacities puts far more
    # output into a block than the bound allows
    cb.predictor_set_params(controls, np.zeros(cb.MLP_PARAMS, np.float32))
    n = 1 << 14
    A, a = pairs_dcsr(controls, *gen.rmat(14, 16, 42), n, n)
    full = oracle.spgemm(A, A)
    bound = max(full.nnz // 5, n)
    plan = cb.dcsr.plan(controls, a, a, mode=cb.SPGEMM_LEARNED, block_bound=bound)
    planned = plan.nblocks
    nb = _check_plan(controls, plan, A, full, bound, "under-predicting plan")
    assert plan.info()["replans"] > 0
    assert nb > planned


@pytest.mark.parametrize("bitmap_pred,combine,window,sparse", [(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 0, 0), (1, 0, 13, 0),
                                                                (0, 0, 14, 0), (1, 0, 15, 0), (1, 0, 0, 1),
                                                                (0, 0, 0, 1 << 20)])
def test_bound_mode_and_bitmap_caps_exact(controls, bitmap_pred, combine, window, sparse):
    # products whose rows reach the bitmap kernels (2^12..2^14 columns) and a
    # wide product (2^21 columns: bitmap rows in several column windows; 2, 4
    # or 8 of them with 2^15, 2^14 or 2^13-word windows); sparse 1 / 2^20:
    # every window slice emitted 64 words per step / by each lane's own words
    base = at.tuning_dict(controls.tuning())
    try:
        at.apply(controls, {**base, "bitmap_pred_caps": bitmap_pred, "gather_combine": combine,
                            "bitmap_window_bits": window, "emit_sparse_bits": sparse})
        p = np.zeros(cb.MLP_PARAMS, np.float32)
        p[-1] = 9.0 / 32.0  # predicts 511 columns for every row
        cb.predictor_set_params(controls, p)
        for scale, wide, seed in ((12, 0, 51), (14, 0, 52), (13, 1 << 21, 53)):
            n = 1 << scale
            r, c = gen.rmat(scale, 16, seed)
            A, a = pairs_dcsr(controls, r, c, n, n)
            if wide:
                cw = (c.astype(np.uint64) * 2654435761 % wide).astype(np.uint32)
                key = np.unique((r.astype(np.uint64) << np.uint64(32)) | cw)
                B, b = pairs_dcsr(controls, (key >> np.uint64(32)).astype(np.uint32),
                                  (key & np.uint64(0xFFFFFFFF)).astype(np.uint32), n, wide)
            else:
                B, b = A, a
            ref = oracle.spgemm(A, B)
            for mode in (cb.SPGEMM_BOUND, cb.SPGEMM_LEARNED):
                got, st = cb.dcsr.spgemm(controls, a, b, mode=mode)
                assert_dcsr_equal(controls, got, ref, f"mode {mode} scale {scale} wide {wide}")
                if mode == cb.SPGEMM_BOUND:
                    assert st["fallback_rows"] == 0
                    assert st["capacity"] == int(np.minimum(oracle.workload(A, B), B.ncols).sum())
    finally:
        at.apply(controls, base)


@pytest.mark.parametrize("mode", [cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED, cb.SPGEMM_BOUND])
def test_run_all_matches_block_runs(controls, mode):
    """Pipelined runs (the next block's passes overlap a one-pass block's
    compaction and emission) give every block exactly, in order; stopping
    early drops the rest."""
    n = 1 << 14
    A, a = pairs_dcsr(controls, *gen.rmat(14, 16, 43), n, n)
    full = oracle.spgemm(A, A)
    cb.predictor_set_params(controls, _over_model())
    bound = max(full.nnz // 6, n)
    seen = []

    def take(k, C, st):
        seen.append((k, C, st))

    plan = cb.dcsr.plan(controls, a, a, mode=mode, block_bound=bound)
    plan.run_all(take)
    assert [k for k, _, _ in seen] == list(range(plan.nblocks)) and len(seen) >= 6
    for k, C, st in seen:
        lo, hi = plan.block(k)
        ref = oracle.spgemm(A, A, row_sel=np.arange(lo, hi, dtype=np.uint32))
        assert_dcsr_equal(controls, C, ref, f"run_all mode {mode} block {k}")
    # the same blocks from single runs
    for k, C, st in seen[:2]:
        D, _ = plan.run(k)
        assert cb.dcsr.digest(controls, D) == cb.dcsr.digest(controls, C)
    # a callback that stops after two blocks
    got = []
    cb.dcsr.plan(controls, a, a, mode=mode, block_bound=bound).run_all(lambda k, C, st: got.append(k) or k == 1)
    assert got == [0, 1]


def test_run_all_recuts_under_pipelining(controls):
    cb.predictor_set_params(controls, np.zeros(cb.MLP_PARAMS, np.float32))
    n = 1 << 14
    A, a = pairs_dcsr(controls, *gen.rmat(14, 16, 44), n, n)
    full = oracle.spgemm(A, A)
    bound = max(full.nnz // 5, n)
    plan = cb.dcsr.plan(controls, a, a, mode=cb.SPGEMM_LEARNED, block_bound=bound)
    blocks = []
    plan.run_all(lambda k, C, st: blocks.append((k, C)))
    assert plan.info()["replans"] > 0 and len(blocks) == plan.nblocks
    rows = 0
    for k, C in blocks:
        lo, hi = plan.block(k)
        assert lo == rows
        ref = oracle.spgemm(A, A, row_sel=np.arange(lo, hi, dtype=np.uint32))
        assert_dcsr_equal(controls, C, ref, f"recut pipelined block {k}")
        assert C.nnz <= bound or hi - lo == 1
        rows = hi
    assert rows == A.nzr
