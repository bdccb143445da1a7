"""The learned row-size predictor and the one-pass SpGEMM.

* The HIP MLP kernels (forward, fused backward, Adam) against the CPU oracle.
* Training on harvested exact row sizes lowers the loss.
* The one-pass SpGEMM is bit-exact against the oracle with a trained model,
  with an under-predicting model (every row re-run) and with an
  over-predicting one.
* What the trained network gains over sizing without it.
"""
import time

import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal, pairs_dcsr

pytestmark = pytest.mark.gpu
F, H1, H2 = 5, 32, 32
# fp32 kernels against a double-precision oracle
RTOL, ATOL = 2e-4, 2e-6


def _rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(cb.MLP_PARAMS) * scale).astype(np.float32)


def test_mlp_forward_matches_oracle(controls):
    p = _rand_params(1)
    x = np.random.default_rng(2).random((1000, F)).astype(np.float32)
    y = cb.mlp_forward(controls, p, x)
    ref = oracle.mlp_forward(p, x, F, H1, H2)
    assert np.allclose(y, ref, rtol=RTOL, atol=1e-5)


@pytest.mark.parametrize("n", [1, 127, 128, 1000, 4096])
def test_mlp_gradient_matches_oracle(controls, n):
    p = _rand_params(3 + n)
    rng = np.random.default_rng(n)
    x = rng.random((n, F)).astype(np.float32)
    t = rng.random(n).astype(np.float32)
    loss, g = cb.mlp_loss_grad(controls, p, x, t)
    rloss, rg = oracle.mlp_loss_grad(p, x, t, F, H1, H2)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss) + 1e-7
    assert np.allclose(g, rg, rtol=1e-3, atol=1e-6 * max(1.0, np.abs(rg).max())), np.abs(g - rg).max()


def test_adam_matches_oracle(controls):
    rng = np.random.default_rng(5)
    p = rng.standard_normal(cb.MLP_PARAMS).astype(np.float32)
    m = (rng.standard_normal(cb.MLP_PARAMS) * 0.1).astype(np.float32)
    v = (rng.random(cb.MLP_PARAMS) * 0.1).astype(np.float32)
    g = rng.standard_normal(cb.MLP_PARAMS).astype(np.float32)
    for step in (1, 7):
        pg, mg, vg = p.copy(), m.copy(), v.copy()
        cb.mlp_adam(controls, pg, mg, vg, g, 1e-3, 0.9, 0.999, 1e-8, step)
        pr, mr, vr = p.copy(), m.copy(), v.copy()
        oracle.adam_step(pr, mr, vr, g, 1e-3, 0.9, 0.999, 1e-8, step)
        assert np.allclose(pg, pr, rtol=1e-6, atol=1e-6)
        assert np.allclose(mg, mr, rtol=1e-6, atol=1e-7) and np.allclose(vg, vr, rtol=1e-5, atol=1e-8)


def _train_set(controls):
    """Harvest exact row sizes from generated power-law and uniform products."""
    cb.predictor_reset(controls, seed=7)
    mats = []
    for k, scale in enumerate((10, 11, 12, 13)):
        n = 1 << scale
        A, a = pairs_dcsr(controls, *gen.rmat(scale, 8 + 4 * k, 40 + k), n, n)
        mats.append((A, a, n))
        cb.predictor_harvest(controls, a, a)
    for seed in (1, 2):
        r, c = gen.uniform_pairs(20000, 2048, 2048, seed)
        A, a = pairs_dcsr(controls, r, c, 2048, 2048)
        cb.predictor_harvest(controls, a, a)
    r, c = gen.hypersparse_zipf(seed=3, log_n=20, row_frac=0.05)
    A, a = pairs_dcsr(controls, r, c, 1 << 20, 1 << 20)
    total = cb.predictor_harvest(controls, a, a)
    return total, mats


@pytest.fixture(scope="module")
def trained(controls):
    total, mats = _train_set(controls)
    st = cb.predictor_train(controls, epochs=40, batch=1024, lr=3e-3, seed=1)
    return total, st, cb.predictor_get_params(controls)


def test_training_lowers_the_loss(controls, trained):
    total, st, _ = trained
    assert total > 5000
    assert st["loss_last"] < 0.25 * st["loss_first"], st
    assert st["loss_last"] < 2e-3, st  # ~ a factor 2^(32*sqrt(2*2e-3)) = 4.2 rms on nnz


def test_learned_spgemm_bit_exact(controls, trained):
    cb.predictor_set_params(controls, trained[2])
    cases = [(gen.rmat(12, 16, 901), 1 << 12), (gen.rmat(14, 16, 902), 1 << 14),
             (gen.uniform_pairs(10000, 1024, 1024, 903), 1024)]
    for (r, c), n in cases:
        A, a = pairs_dcsr(controls, r, c, n, n)
        got, st = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED)
        assert_dcsr_equal(controls, got, oracle.spgemm(A, A), f"learned n={n}")
        assert st["products"] == int(oracle.workload(A, A).sum())
        assert st["fallback_rows"] <= 0.2 * A.nzr


def test_learned_row_blocks_and_hypersparse(controls, trained):
    cb.predictor_set_params(controls, trained[2])
    r, c = gen.hypersparse_zipf(seed=11, log_n=24, row_frac=0.01)
    n = 1 << 24
    A, a = pairs_dcsr(controls, r, c, n, n)
    got, st = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED)
    assert_dcsr_equal(controls, got, oracle.spgemm(A, A), "learned hypersparse")
    r, c = gen.rmat(13, 16, 12)
    A, a = pairs_dcsr(controls, r, c, 1 << 13, 1 << 13)
    full = oracle.spgemm(A, A)
    lo, hi = 100, A.nzr // 2
    blk, _ = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED, row_begin=lo, row_end=hi)
    ref = oracle.spgemm(A, A, row_sel=np.arange(lo, hi, dtype=np.uint32))
    assert_dcsr_equal(controls, blk, ref, "learned block")


def test_underpredicting_model_reruns_every_row_exactly(controls):
    # all weights zero, output bias zero: predicts 0 for every row, so every
    # hash row overflows its capacity and is re-run exactly
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    cb.predictor_set_params(controls, p)
    # 2^14 columns: a 16384-bit bitmap is larger than the small hash tables
    # the zero prediction asks for, so rows go to hash bins and overflow
    r, c = gen.rmat(14, 16, 21)
    A, a = pairs_dcsr(controls, r, c, 1 << 14, 1 << 14)
    got, st = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED)
    ref = oracle.spgemm(A, A)
    assert_dcsr_equal(controls, got, ref, "under-predicting")
    # rows that can overflow: more distinct columns than the padded capacity,
    # outside the tiny and bitmap kernels (which get the exact upper bound)
    lens = np.diff(ref.rpt.astype(np.int64))
    assert st["fallback_rows"] > 0
    assert st["fallback_rows"] <= int((lens > 8).sum())


def test_midsize_underprediction_fails_fast_and_exact(controls):
    """A model that predicts 2047 columns for every row puts rows with up to
    a million distinct columns (over 2^22) into 2^13-slot LDS hash tables.
    A full table must fail its row after a bounded probe sequence -- not probe
    the whole table once per product -- and the failed rows are re-run
    exactly."""
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    p[-1] = 11.0 / 32.0  # log2(1 + N) / 32 = 11/32: N = 2047
    cb.predictor_set_params(controls, p)
    rng = np.random.default_rng(31)
    hubs, ncols = 64, 1 << 22
    lens = rng.choice([300, 3000, 12000, 40000], hubs)
    rb = np.repeat(np.arange(hubs, dtype=np.uint32), lens)
    cbv = rng.integers(0, ncols, rb.size).astype(np.uint32)
    k = rng.integers(1, 31, 300)
    ra = np.repeat(np.arange(300, dtype=np.uint32), k)
    ca = rng.integers(0, hubs, ra.size).astype(np.uint32)
    A, a = pairs_dcsr(controls, *np.unique(np.c_[ra, ca], axis=0).T, 300, hubs)
    B, b = pairs_dcsr(controls, *np.unique(np.c_[rb, cbv], axis=0).T, hubs, ncols)
    t0 = time.perf_counter()
    got, st = cb.dcsr.spgemm(controls, a, b, mode=cb.SPGEMM_LEARNED)
    controls.sync()
    elapsed = time.perf_counter() - t0
    assert_dcsr_equal(controls, got, oracle.spgemm(A, B), "mid-size under-prediction")
    assert st["fallback_rows"] > 0
    assert elapsed < 10.0, f"one pass with full hash tables took {elapsed:.1f} s"


def test_overpredicting_model_is_exact(controls):
    p = np.zeros(cb.MLP_PARAMS, np.float32)
    p[-1] = 1.0  # predicts 2^32 - 1: every capacity clamps to min(P, ncols)
    cb.predictor_set_params(controls, p)
    r, c = gen.rmat(12, 16, 22)
    A, a = pairs_dcsr(controls, r, c, 1 << 12, 1 << 12)
    got, st = cb.dcsr.spgemm(controls, a, a, mode=cb.SPGEMM_LEARNED)
    assert_dcsr_equal(controls, got, oracle.spgemm(A, A), "over-predicting")
    assert st["fallback_rows"] == 0
    assert st["capacity"] == int(np.minimum(oracle.workload(A, A), 1 << 12).sum())


def test_trained_network_gains_over_sizing_without_it(controls, trained):
    """Without the network a single pass has two choices: size every row by
    the only safe bound min(P, ncols) (the over-predicting model), or guess
    small and re-run what overflows (the zero model).  On a product whose
    columns span 2^22 (so LDS hash tables, not bitmaps, hold most rows) the
    trained model reserves far less memory than the bound, re-runs almost no
    rows, and is faster than re-running them."""
    n, wide = 1 << 15, 1 << 22
    r, c = gen.rmat(15, 16, 31)
    A, a = pairs_dcsr(controls, r, c, n, n)
    cw = (c.astype(np.uint64) * 2654435761 % wide).astype(np.uint32)
    key = np.unique((r.astype(np.uint64) << np.uint64(32)) | cw)
    B, b = pairs_dcsr(controls, (key >> np.uint64(32)).astype(np.uint32),
                      (key & np.uint64(0xFFFFFFFF)).astype(np.uint32), n, wide)
    ref = oracle.spgemm(A, B)
    zero = np.zeros(cb.MLP_PARAMS, np.float32)
    over = zero.copy()
    over[-1] = 1.0
    res = {}
    for name, params in (("trained", trained[2]), ("zero", zero), ("bound", over)):
        cb.predictor_set_params(controls, params)
        got, st = cb.dcsr.spgemm(controls, a, b, mode=cb.SPGEMM_LEARNED)
        assert_dcsr_equal(controls, got, ref, name)
        best = 1e9
        for _ in range(3):
            t0 = time.perf_counter()
            cb.dcsr.spgemm(controls, a, b, mode=cb.SPGEMM_LEARNED)
            best = min(best, time.perf_counter() - t0)
        res[name] = {"capacity": st["capacity"], "fallback": st["fallback_rows"], "ms": 1e3 * best,
                     "err": st["pred_abs_log2_err"]}
    print(res, "nnz", ref.nnz, "rows", A.nzr)
    t, z, bnd = res["trained"], res["zero"], res["bound"]
    assert t["fallback"] < 0.05 * A.nzr and t["fallback"] * 5 < z["fallback"]
    assert bnd["fallback"] == 0
    # every row reserves its prediction (with the margin), far below the bound
    assert 0.8 * ref.nnz <= t["capacity"] < bnd["capacity"]
    assert t["err"] < 0.5
    assert t["ms"] < z["ms"]
