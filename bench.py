"""Benchmark of the boolean SpGEMM hot path on one MI355X.

Workload (default): power-law A * A with A an R-MAT graph of scale 22, edge
factor 16, (a, b, c, d) = (.57, .19, .19, .05), seed 1, duplicates and
self-loops removed, uint32 DCSR; the same at scale 20 is reported beside it.
Their products hold about 7e10 and 1e10 nonzeros, more than uint32 indices
hold, so a step multiplies A by A in row blocks whose results each fit.

A "step" is one full A * A with the inputs resident in HBM, in each mode:
plain (two passes), learned (one pass sized by the predictor) and bound (one
pass sized by min(products, ncols), no predictor).  Each mode plans its own
row blocks inside the step from what it knows (exact sizes from its symbolic
pass, predicted capacities, or bounds; clb_spgemm_plan_*).  `value` is
intermediate products per second of the learned step (the sum over rows of
sum_{k in A_i} |B_k|, as count_workload.cl:47-51 counts them).  The JSON line
also carries the roofline of the dominant kernel (algorithmic bytes per
launch over its HIP-event duration) and the CPU baseline (the reference test
suite's host SpGEMM, matrix_multiplication_cpu restated in oracle/, timed on
a bounded row sample) beside a multi-threaded Gustavson on the host cores.
Reported beside the main metric as well: the 64-label closure (4 steps; to
its fixed point with --closure-only), the hypersparse learned-sizing workload
(2^24 x 2^24, the predictor trained on 200 instances) and the automaton (x)
R-MAT 22/32 Kronecker with one K * K step at G scale 16.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload rmat22] [--also rmat20]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "intermediate products/s, learned+batched vs plain hash SpGEMM, bit-exact"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (scale, edge_factor, seed)
    "rmat20": (20, 16, 1),
    "rmat22": (22, 16, 1),
    "rmat18": (18, 16, 1),
    "rmat16": (16, 16, 1),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def row_products(rpt, rows, cols, nrows_b, b_rpt, b_rows):
    blen = np.zeros(nrows_b, np.int64)
    blen[b_rows] = np.diff(b_rpt.astype(np.int64))
    pe = blen[cols]
    return np.add.reduceat(pe, rpt[:-1].astype(np.int64)) if rows.size else np.zeros(0, np.int64)


def row_blocks(prod, budget, ncols=None):
    """Split stored rows into consecutive blocks whose per-row products add up
    to at most `budget`; with `ncols`, each row counts min(products, ncols),
    an upper bound of its output size (a budget below 2^32 then guarantees
    that every block's result fits uint32 indices)."""
    w = np.minimum(prod, ncols) if ncols else prod
    bounds = [0]
    acc = 0
    for i, p in enumerate(w.tolist()):
        if acc + p > budget and i > bounds[-1]:
            bounds.append(i)
            acc = 0
        acc += p
    bounds.append(len(prod))
    return list(zip(bounds[:-1], bounds[1:]))


def validate_sample(cb, oracle, ctl, a_dev, A, blocks, mode, rng, rows=480):
    """Bit-exact check of sampled rows of every block against the oracle
    (about `rows` rows in all, at least 48 per block)."""
    rows_per_block = max(48, -(-rows // max(1, len(blocks))))
    checked = 0
    for lo, hi in blocks:
        C, st = cb.dcsr.spgemm(ctl, a_dev, a_dev, mode=mode, row_begin=lo, row_end=hi)
        sel = np.unique(rng.integers(lo, hi, min(rows_per_block, hi - lo))).astype(np.uint32)
        ref = oracle.spgemm(A, A, row_sel=sel)
        rpt = C.rpt_gpu.read(ctl, C.nzr + 1)
        crow = C.rows_gpu.read(ctl, C.nzr)
        pos = np.searchsorted(crow, ref.rows)
        for q in range(ref.nzr):
            i = int(pos[q])
            if i >= crow.size or crow[i] != ref.rows[q]:
                raise AssertionError(f"row {ref.rows[q]} missing in block {lo}-{hi}")
            n = int(rpt[i + 1] - rpt[i])
            got = C.cols_gpu.read(ctl, n, offset=4 * int(rpt[i]))
            if not np.array_equal(got, ref.cols[ref.rpt[q]:ref.rpt[q + 1]]):
                raise AssertionError(f"row {ref.rows[q]} differs in block {lo}-{hi}")
            checked += 1
        # a sampled row is stored on the GPU exactly when the oracle finds it nonempty
        gpu_rows, ref_rows = set(crow.tolist()), set(ref.rows.tolist())
        for r in A.rows[sel].tolist():
            if (r in gpu_rows) != (r in ref_rows):
                raise AssertionError(f"row {r}: stored on GPU={r in gpu_rows}, oracle nonempty={r in ref_rows}")
        del C
    return checked


def _new_agg():
    return {"products": 0, "nnz": 0, "ms_kernel": [0.0] * 16, "launches": [0] * 16,
            "bin_rows": [0] * 32, "bin_products": [0] * 32, "bin_a_nnz": [0] * 32, "bin_out_nnz": [0] * 32,
            "fallback_rows": 0, "dense_rows": 0, "dense_nnz": 0, "blocks": 0, "replans": 0, "plan_ms": 0.0,
            "err_rows": 0, "err_sum": 0.0, "capacity": 0}


def _add_stats(agg, st):
    agg["products"] += st["products"]
    agg["nnz"] += st["nnz"]
    agg["fallback_rows"] += st["fallback_rows"]
    agg["dense_rows"] += st["dense_rows"]
    agg["dense_nnz"] += st["dense_nnz"]
    agg["capacity"] += st["capacity"]
    agg["err_rows"] += st["nzr"]
    agg["err_sum"] += st["pred_abs_log2_err"] * st["nzr"]
    for k in ("ms_kernel", "launches", "bin_rows", "bin_products", "bin_a_nnz", "bin_out_nnz"):
        agg[k] = [x + y for x, y in zip(agg[k], st[k])]


def run_steps(cb, ctl, a_dev, mode, n, timing, block_bound):
    """n full A*A products.  Each step plans its own row blocks (the mode's
    sizing pass over all rows of A: symbolic counts, predicted capacities or
    bounds) and then runs every block (clb_spgemm_plan_run_all); the planning
    is inside the timed region.  Returns (seconds, statistics of the last
    step)."""
    agg = None
    t0 = time.perf_counter()
    for _ in range(n):
        agg = _new_agg()
        plan = cb.dcsr.plan(ctl, a_dev, a_dev, mode=mode, block_bound=block_bound)
        agg["plan_ms"] = plan.info()["ms_plan"]
        # blocks run back to back (one-pass blocks overlap); each result is
        # dropped as soon as it is complete
        plan.run_all(lambda k, C, st: _add_stats(agg, st), timing=timing)
        info = plan.info()
        agg["blocks"], agg["replans"] = info["nblocks"], info["replans"]
        plan.close()
    ctl.sync()
    return time.perf_counter() - t0, agg


def plan_blocks(cb, ctl, a_dev, mode, block_bound):
    """The row blocks a plan of this mode cuts (for validation)."""
    plan = cb.dcsr.plan(ctl, a_dev, a_dev, mode=mode, block_bound=block_bound)
    blocks = [plan.block(k) for k in range(plan.nblocks)]
    plan.close()
    return blocks


def train_predictor(cb, gen, ctl, log):
    """Train the learned row-size predictor on the GPU from exact row sizes of
    other generated instances (never the benchmark matrix itself)."""
    t0 = time.perf_counter()
    cb.predictor_reset(ctl, seed=1)
    total = 0
    recipe = [("rmat", 16, 16, 2), ("rmat", 17, 16, 3), ("rmat", 18, 16, 4), ("rmat", 17, 8, 5), ("rmat", 16, 32, 6),
              ("uniform", 1 << 14, 20, 7), ("hyper", 22, 0, 8)]
    for kind, x, y, seed in recipe:
        if kind == "rmat":
            r, c = gen.rmat(x, y, seed)
            n = 1 << x
        elif kind == "uniform":
            r, c = gen.uniform_pairs(x * y, x, x, seed)
            n = x
        else:
            r, c = gen.hypersparse_zipf(seed=seed, log_n=x, row_frac=0.01)
            n = 1 << x
        rpt, rows, cols = gen.to_dcsr(r, c)
        m = cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n, n)
        total = cb.predictor_harvest(ctl, m, m, max_rows=1 << 17)
    st = cb.predictor_train(ctl, epochs=30, batch=4096, lr=3e-3, seed=1)
    st["samples"] = total
    st["wall_s"] = time.perf_counter() - t0
    log(f"[bench] predictor trained on {total} rows: loss {st['loss_first']:.5f} -> {st['loss_last']:.5f} "
        f"({st['ms']:.0f} ms of training)")
    return st


def _fingerprint(ctl, m):
    """(nnz, nzr, device-side digest of the whole matrix): equal fingerprints
    of two results mean identical matrices (clb_dcsr_digest)."""
    import clbool_amd as cb
    return (int(m.nnz), int(m.nzr), cb.dcsr.digest(ctl, m))


def bench_closure(cb, gen, ctl, max_iters, log, seed=0, legs=None, warm=True):
    """The 64-label closure on 2^18 vertices: 32 seeded triples, X |= Y * Z to
    a fixed point or max_iters steps.  Per-triple plain SpGEMMs (the
    reference's call pattern), batched plain SpGEMMs (batching alone) and
    batched learned SpGEMMs; all runs start from the same matrices and must
    end with identical ones.
    Each mode starts from an emptied buffer pool and runs the whole closure
    once untimed, so both are timed with their own buffers cached."""
    n = 1 << 18
    labs = gen.closure_labels(n_labels=64, log_n=18, seed=seed)
    triples = gen.closure_triples(64, 32, seed=seed + 1)
    dcsr = [gen.to_dcsr(r, c) for r, c in labs]
    out = {"labels": 64, "vertices": n, "triples": 32, "label_nnz": int(sum(d[2].size for d in dcsr))}
    finals = {}
    all_legs = (("per_triple_plain", cb.CLOSURE_PER_TRIPLE, cb.SPGEMM_PLAIN),
                ("batched_plain", cb.CLOSURE_BATCHED, cb.SPGEMM_PLAIN),
                ("batched_learned", cb.CLOSURE_BATCHED, cb.SPGEMM_LEARNED))
    for name, mode, smode in [x for x in all_legs if legs is None or x[0] in legs]:
        ctl.trim()
        if warm:
            mats = [cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n) for rpt, rows, cols in dcsr]
            for it in range(max_iters):  # warm-up
                if cb.closure_step(ctl, mats, triples, mode=mode, spgemm_mode=smode)["changed"] == 0:
                    break
            del mats
        mats = [cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n) for rpt, rows, cols in dcsr]
        ctl.sync()
        products, iters, steps_ms, step_nnz = 0, 0, [], []
        t0 = time.perf_counter()
        step_mallocs, step_phases = [], []
        for it in range(max_iters):
            ps0 = ctl.pool_stats()
            ts = time.perf_counter()
            log(f"[bench] closure {name} step {it + 1} starts at monotonic {time.monotonic_ns()} ns")
            st = cb.closure_step(ctl, mats, triples, mode=mode, spgemm_mode=smode)
            steps_ms.append(1e3 * (time.perf_counter() - ts))
            ps1 = ctl.pool_stats()
            step_mallocs.append([ps1["mallocs"] - ps0["mallocs"], round((ps1["malloc_bytes"] - ps0["malloc_bytes"]) / 1e9, 2),
                                 ps1["drops"] - ps0["drops"], round(ps0["free_bytes"] / 1e9, 1),
                                 round(ps0["cached_bytes"] / 1e9, 1)])
            step_phases.append([round(st["ms_products"], 1), round(st["ms_unions"], 1), st["fallback_rows"],
                                st["dense_rows"]])
            step_nnz.append(int(sum(m.nnz for m in mats)))
            products += st["products"]
            iters += 1
            if st["changed"] == 0:
                break
        secs = time.perf_counter() - t0
        # every matrix compared in full by its device-side digest
        finals[name] = [cb.dcsr.digest(ctl, m) for m in mats]
        out[name] = {"iterations": iters, "seconds": secs, "products": products,
                     "products_per_s": products / secs, "step_ms": [round(x, 2) for x in steps_ms],
                     "step_nnz": step_nnz, "fixed_point": st["changed"] == 0, "final_nnz": step_nnz[-1],
                     # per step: hipMalloc calls, GB they allocated, cache drops, free and pool-cached
                     # device GB at the step's start
                     "step_pool": step_mallocs,
                     # per step: ms of the SpGEMMs, ms of the unions, fallback rows, dumped dense rows
                     "step_phases": step_phases}
        log(f"[bench] closure {name}: {iters} steps, {secs * 1e3:.1f} ms, {products / secs:.3e} products/s")
    # full-matrix digests of all 64 labels
    out["identical_results"] = len({tuple(v) for v in finals.values()}) == 1
    out["warm"] = warm
    sec = {k: out[k]["seconds"] for k in finals}
    if "per_triple_plain" in sec and "batched_learned" in sec:
        out["speedup"] = sec["per_triple_plain"] / sec["batched_learned"]  # learned + batched vs plain
    if "per_triple_plain" in sec and "batched_plain" in sec:
        out["batching_speedup"] = sec["per_triple_plain"] / sec["batched_plain"]  # same SpGEMM mode
    if "batched_plain" in sec and "batched_learned" in sec:
        out["learned_speedup_batched"] = sec["batched_plain"] / sec["batched_learned"]
    out["fixed_point"] = all(out[k]["fixed_point"] for k in finals)
    return out


def _dcsr_rows(rpt, rows, cols, keep):
    """Host DCSR restricted to the stored rows whose positions are in `keep`
    (ascending); row ids unchanged."""
    rpt = rpt.astype(np.int64)
    lens = rpt[keep + 1] - rpt[keep]
    starts = np.repeat(rpt[keep] - np.r_[0, np.cumsum(lens)[:-1]], lens)
    idx = starts + np.arange(int(lens.sum()), dtype=np.int64)
    return np.r_[0, np.cumsum(lens)].astype(np.uint32), rows[keep].astype(np.uint32), cols[idx].astype(np.uint32)


def bench_c5(cb, gen, ctl, log, scale=22, ef=32, step_scale=16, seed=1, tunings=None):
    """The RPQ input of a 64x64 automaton M (200 uniform nonzeros) and an
    R-MAT graph G (scale 22, edge factor 32).  K = M (x) G has about 2.6e10
    nonzeros, more than one uint32 DCSR holds, so it is built in blocks of
    automaton rows (K rows i * 2^scale + g) and timed as a write-bound kernel.
    One closure step K * K at this scale is products(M*M) * products(G*G),
    about 3e14 products, beyond one GPU; the step itself is measured on the
    same automaton with G at scale `step_scale` (K of width 64 * 2^16 = 2^22),
    whole, in row blocks of output bound below 2^32, plain and learned; both
    results must agree.  tunings: knobs per mode ("plain", "learned") applied
    before that mode's step (None: the controls' current knobs for both)."""
    mr, mc = gen.uniform_pairs(200, 64, 64, seed)
    mrpt, mrows, mcols = gen.to_dcsr(mr, mc)
    out = {"automaton": {"n": 64, "nnz": int(mr.size), "seed": seed}}
    t0 = time.perf_counter()
    r, c = gen.rmat(scale, ef, seed)
    grpt, grows, gcols = gen.to_dcsr(r, c)
    del r, c
    n = 1 << scale
    log(f"[bench] C5: G nnz={gcols.size} generated in {time.perf_counter() - t0:.1f} s")
    G = cb.matrix_dcsr_from_cpu(ctl, grpt, grows, gcols, n, n)
    g_products = int(row_products(grpt, grows, gcols, n, grpt, grows).sum())
    out["graph"] = {"scale": scale, "edge_factor": ef, "seed": seed, "nnz": int(gcols.size), "nzr": int(grows.size),
                    "products_g_times_g": g_products}
    del grpt, grows, gcols
    # Kronecker, one automaton row at a time (each block < 2^32 nonzeros)
    kron_nnz, kron_bytes, kron_s = 0, 0, 0.0
    for q in range(mrows.size):
        M1 = cb.matrix_dcsr_from_cpu(ctl, *_dcsr_rows(mrpt, mrows, mcols, np.array([q])), 64, 64)
        ctl.sync()
        t = time.perf_counter()
        K = cb.dcsr.kronecker_product(ctl, M1, G)
        ctl.sync()
        kron_s += time.perf_counter() - t
        kron_nnz += K.nnz
        kron_bytes += 4 * K.nnz + 8 * K.nzr + 4
        del K
    del G
    out["kronecker"] = {"blocks": int(mrows.size), "nnz": kron_nnz, "bytes_written": kron_bytes,
                        "seconds": kron_s, "GB_per_s": kron_bytes / kron_s / 1e9,
                        "frac_of_hbm_peak": kron_bytes / kron_s / 1e9 / HBM_PEAK_GBS,
                        "note": "host-timed per block (includes the size readback and pooled allocation)"}
    log(f"[bench] C5 Kronecker: {kron_nnz} nnz in {mrows.size} blocks, {kron_s * 1e3:.1f} ms, "
        f"{kron_bytes / kron_s / 1e9:.0f} GB/s written")
    # one closure step K * K with G at step_scale
    mlen = np.zeros(64, np.int64)
    mlen[mrows] = np.diff(mrpt.astype(np.int64))
    two_step = np.add.reduceat(mlen[mcols], mrpt[:-1].astype(np.int64))  # sum over j in M_i of |M_j|
    r, c = gen.rmat(step_scale, ef, seed)
    hrpt, hrows, hcols = gen.to_dcsr(r, c)
    m = 1 << step_scale
    pg = row_products(hrpt, hrows, hcols, m, hrpt, hrows)
    H = cb.matrix_dcsr_from_cpu(ctl, hrpt, hrows, hcols, m, m)
    Mall = cb.matrix_dcsr_from_cpu(ctl, mrpt, mrows, mcols, 64, 64)
    K = cb.dcsr.kronecker_product(ctl, Mall, H)
    # K's stored rows are (i, g) in that order; row (i, g) has
    # sum_{j in M_i} sum_{h in G_g} |M_j| |G_h| = two_step[i] * products(G_g) products
    prod = np.concatenate([int(s) * pg for s in two_step])
    blocks = row_blocks(prod, float((1 << 32) - 1), ncols=64 * m)
    step = {"graph_scale": step_scale, "k_nrows": 64 * m, "k_nnz": int(K.nnz), "products": int(prod.sum()),
            "row_blocks": len(blocks), "step_products_at_full_scale": int(two_step.sum()) * g_products}
    train = train_predictor(cb, gen, ctl, log)
    prints = {}
    for name, mode in (("plain", cb.SPGEMM_PLAIN), ("learned", cb.SPGEMM_LEARNED)):
        if tunings and name in tunings:
            from clbool_amd import autotune as at
            at.apply(ctl, tunings[name])
        best = None
        for rep in range(3):
            ctl.sync()
            t = time.perf_counter()
            nnz, fp = 0, []
            for lo, hi in blocks:
                Cb, st = cb.dcsr.spgemm(ctl, K, K, mode=mode, row_begin=lo, row_end=hi)
                nnz += Cb.nnz
                if rep == 2:
                    fp.append(_fingerprint(ctl, Cb))
                del Cb
            ctl.sync()
            dt = time.perf_counter() - t
            if rep > 0:
                best = dt if best is None else min(best, dt)
        prints[name] = fp
        step[name] = {"seconds": best, "products_per_s": float(prod.sum()) / best, "nnz": nnz}
        log(f"[bench] C5 K*K step (G scale {step_scale}) {name}: {prod.sum():.3e} products in {best * 1e3:.1f} ms, "
            f"{prod.sum() / best:.3e} products/s, nnz {nnz}")
    step["identical_results"] = prints["plain"] == prints["learned"]
    step["learned_vs_plain"] = step["plain"]["seconds"] / step["learned"]["seconds"]
    step["predictor_training_ms"] = train["ms"]
    out["kk_step"] = step
    return out


def bench_hypersparse(cb, gen, ctl, n_train, log, n_test=4, log_n=24):
    """The hypersparse learned-sizing workload: 2^24 x 2^24, 1% of rows
    nonempty, Zipf(1.5) row lengths capped at 4096, uniform columns; A * A.
    The predictor is trained on exact row sizes of n_train seeded instances
    (none of them a test instance), then plain and learned SpGEMMs of n_test
    other instances are compared: identical results, fallback rate, and the
    error of the predicted row sizes."""
    n = 1 << log_n
    t0 = time.perf_counter()
    cb.predictor_reset(ctl, seed=1)
    total, nnz_seen = 0, []
    for seed in range(1, n_train + 1):
        r, c = gen.hypersparse_zipf(seed=seed, log_n=log_n, row_frac=0.01)
        rpt, rows, cols = gen.to_dcsr(r, c)
        nnz_seen.append(int(cols.size))
        m = cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n, n)
        total = cb.predictor_harvest(ctl, m, m, max_rows=1 << 24)
        del m
    gen_s = time.perf_counter() - t0
    tr = cb.predictor_train(ctl, epochs=20, batch=4096, lr=3e-3, seed=1)
    log(f"[bench] hypersparse: {n_train} instances ({np.mean(nnz_seen):.3g} nnz each), {total} samples, "
        f"loss {tr['loss_first']:.5f} -> {tr['loss_last']:.6f}")
    tests = []
    for seed in range(100001, 100001 + n_test):
        r, c = gen.hypersparse_zipf(seed=seed, log_n=log_n, row_frac=0.01)
        rpt, rows, cols = gen.to_dcsr(r, c)
        a = cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n, n)
        res = {}
        for name, mode in (("plain", cb.SPGEMM_PLAIN), ("learned", cb.SPGEMM_LEARNED)):
            C, st = cb.dcsr.spgemm(ctl, a, a, mode=mode)  # warm
            ts = []
            for _ in range(5):
                t1 = time.perf_counter()
                C, st = cb.dcsr.spgemm(ctl, a, a, mode=mode)
                ctl.sync()
                ts.append(time.perf_counter() - t1)
            res[name] = (C, st, min(ts))
        Cp, stp, tp = res["plain"]
        Cl, stl, tl = res["learned"]
        same = _fingerprint(ctl, Cp) == _fingerprint(ctl, Cl)
        # predicted against exact row sizes, per stored row of A
        pred = cb.predictor_predict(ctl, a, a).astype(np.float64)
        exact = np.zeros(rows.size, np.float64)
        if Cp.nnz:
            crow = Cp.rows_gpu.read(ctl, Cp.nzr)
            crpt = Cp.rpt_gpu.read(ctl, Cp.nzr + 1).astype(np.int64)
            exact[np.searchsorted(rows, crow)] = np.diff(crpt)
        nz = exact > 0
        tests.append({"seed": seed, "a_nnz": int(cols.size), "a_nzr": int(rows.size), "products": int(stp["products"]),
                      "c_nnz": int(Cp.nnz), "identical": bool(same),
                      "fallback_rows": int(stl["fallback_rows"]), "fallback_rate": stl["fallback_rows"] / max(1, rows.size),
                      "mae_rows": float(np.abs(pred - exact).mean()),
                      "mape_rows": float((np.abs(pred - exact)[nz] / exact[nz]).mean()) if nz.any() else 0.0,
                      "mean_abs_log2_err": float(np.abs(np.log2(1 + np.maximum(pred, 0)) - np.log2(1 + exact)).mean()),
                      "capacity_over_nnz": stl["capacity"] / max(1, Cp.nnz),
                      "plain_ms": tp * 1e3, "learned_ms": tl * 1e3})
        log(f"[bench] hypersparse test {seed}: identical={same} fallback={stl['fallback_rows']} "
            f"mae={tests[-1]['mae_rows']:.3f} plain {tp * 1e3:.2f} ms learned {tl * 1e3:.2f} ms")
        del a, Cp, Cl
    return {"train_instances": n_train, "train_samples": int(total), "train_nnz_mean": float(np.mean(nnz_seen)),
            "generation_and_harvest_s": gen_s, "training": tr, "tests": tests}


def cpu_baseline(oracle, A, target_s=12.0, seed=7):
    """matrix_multiplication_cpu (coo_utils.cpp:470-514) restated in oracle/,
    single thread, on a bounded uniform sample of A's rows times the full B."""
    rng = np.random.default_rng(seed)
    n_rows = 16
    spent, rows_done, products = 0.0, 0, 0
    while spent < target_s and rows_done < A.nzr:
        sel = np.sort(rng.choice(A.nzr, min(n_rows, A.nzr), replace=False))
        sub_rpt = np.r_[0, np.cumsum(np.diff(A.rpt.astype(np.int64))[sel])].astype(np.uint32)
        sub_cols = np.concatenate([A.cols[A.rpt[i]:A.rpt[i + 1]] for i in sel]).astype(np.uint32)
        S = oracle.Dcsr(sub_rpt, A.rows[sel].astype(np.uint32), sub_cols, A.nrows, A.ncols)
        t0 = time.perf_counter()
        oracle.spgemm_reference(S, A)
        dt = time.perf_counter() - t0
        spent += dt
        rows_done += sel.size
        products += int(oracle.workload(S, A).sum())
        n_rows = max(16, int(n_rows * min(4.0, max(1.5, (target_s - spent) / max(dt, 1e-3)))))
    return {"value": products / spent if spent else None, "unit": "products/s", "cores": 1, "kind": "port",
            "sample": f"{rows_done} uniformly sampled A rows (seed {seed}) times the full B, "
                      f"{products} products in {spent:.1f} s"}


def cpu_gustavson(oracle, A, threads, target_rows=20000, seed=11):
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(A.nzr, min(target_rows, A.nzr), replace=False)).astype(np.uint32)
    t0 = time.perf_counter()
    oracle.spgemm(A, A, row_sel=sel, threads=threads)
    dt = time.perf_counter() - t0
    S = oracle.Dcsr(np.r_[0, np.cumsum(np.diff(A.rpt.astype(np.int64))[sel])].astype(np.uint32), A.rows[sel],
                    np.concatenate([A.cols[A.rpt[i]:A.rpt[i + 1]] for i in sel]).astype(np.uint32))
    p = int(oracle.workload(S, A).sum())
    return {"value": p / dt, "unit": "products/s", "cores": threads,
            "sample": f"{sel.size} sampled rows, {p} products in {dt:.2f} s (Gustavson, dense marker)"}


MODE_CODES = {"plain": 0, "learned": 1, "bound": 2}


def roofline_of(cb, agg, workload):
    """Roofline of the dominant kernel family of a timed step: algorithmic
    bytes of that family's launches over their HIP-event durations."""
    fam_names = cb._capi.KERNEL_FAMILIES
    fam = int(np.argmax(agg["ms_kernel"]))
    launches = max(agg["launches"][fam], 1)
    avg_ms = agg["ms_kernel"][fam] / launches
    # A rows (rpt 8 B/row, a_cols 4 B + B row extent 8 B per A nonzero), 4 B
    # per product read from B, 4 B per output column the family itself
    # writes.  Dense bitmap rows are written by the emission family (4 B per
    # column), not by the bitmap family that found them; the bitmap dump in
    # between is not counted.  The compaction moves 4 B in and 4 B out per
    # column it copies.
    pass_, kind = divmod(fam, 4)
    bins = {0: range(1, 2), 1: range(2, 13), 2: range(13, 32)}.get(kind, range(0))
    writes = pass_ != 0
    rows_f = sum(agg["bin_rows"][b] for b in bins)
    bytes_total = sum(12 * agg["bin_a_nnz"][b] + 4 * agg["bin_products"][b] +
                      (4 * agg["bin_out_nnz"][b] if writes else 0) for b in bins) + 8 * rows_f
    if kind == 2 and writes:
        bytes_total -= 4 * agg["dense_nnz"]
    if fam_names[fam] == "onepass_compact":
        bytes_total = 8 * (agg["nnz"] - agg["dense_nnz"])
    elif kind == 3:
        bytes_total = 4 * agg["dense_nnz"]
    bytes_per_launch = bytes_total / launches
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(fam_names[fam])
        except Exception:
            traffic = None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": fam_names[fam], "launches": launches, "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_bytes_per_launch": int(bytes_per_launch)}


def step_algorithmic_bytes(a_nzr, a_nnz, products, c_nzr_est, nnz):
    """Compulsory bytes of one A*B over all blocks (single pass over every
    operand and the output)."""
    return 4 * (2 * a_nzr + 1) + 12 * a_nnz + 4 * products + 4 * (2 * c_nzr_est + 1) + 4 * nnz


def bench_product(cb, gen, ctl, oracle, workload, modes, steps, warmup, tuned, validate, block_bound, log):
    """A * A of an R-MAT workload, every mode planning its own blocks inside
    the timed step."""
    from clbool_amd import autotune as at
    scale, ef, seed = WORKLOADS[workload]
    t0 = time.perf_counter()
    r, c = gen.rmat(scale, ef, seed)
    rpt, rows, cols = gen.to_dcsr(r, c)
    n = 1 << scale
    del r, c
    log(f"[bench] {workload}: R-MAT scale {scale} ef {ef} seed {seed} generated in {time.perf_counter() - t0:.1f} s")
    a_dev = cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, n, n)
    prod = row_products(rpt, rows, cols, n, rpt, rows)
    total_products = int(prod.sum())
    A = oracle.Dcsr(rpt, rows, cols, n, n)
    config = {"workload": f"rmat{scale}_ef{ef}_AxA", "scale": scale, "edge_factor": ef, "seed": seed,
              "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "a_nnz": int(cols.size), "a_nzr": int(rows.size),
              "products": total_products, "block_bound": int(block_bound)}
    results, validated = {}, {}
    for m in modes:
        at.apply(ctl, tuned[m])
        if validate:
            blocks = plan_blocks(cb, ctl, a_dev, MODE_CODES[m], block_bound)
            validated[m] = validate_sample(cb, oracle, ctl, a_dev, A, blocks, MODE_CODES[m],
                                           np.random.default_rng(3))
            log(f"[bench] {workload} {m}: {validated[m]} sampled rows bit-exact vs oracle")
        ctl.trim()
        run_steps(cb, ctl, a_dev, MODE_CODES[m], warmup, False, block_bound)
        ps0 = ctl.pool_stats()
        secs, agg = run_steps(cb, ctl, a_dev, MODE_CODES[m], steps, True, block_bound)
        ps1 = ctl.pool_stats()
        agg["pool_timed"] = {"mallocs": ps1["mallocs"] - ps0["mallocs"],
                             "malloc_GB": (ps1["malloc_bytes"] - ps0["malloc_bytes"]) / 1e9,
                             "cache_drops": ps1["drops"] - ps0["drops"]}
        ms = 1e3 * secs / steps
        results[m] = {"ms_per_step": ms, "products_per_s": agg["products"] / (ms / 1e3),
                      "nnz_per_s": agg["nnz"] / (ms / 1e3), "nnz": agg["nnz"], "agg": agg}
        log(f"[bench] {workload} {m}: {ms:.2f} ms/step ({agg['blocks']} blocks, plan {agg['plan_ms']:.1f} ms), "
            f"{results[m]['products_per_s']:.3e} products/s, fallback rows {agg['fallback_rows']}")
        assert agg["products"] == total_products
    nnzs = {m: v["nnz"] for m, v in results.items()}
    assert len(set(nnzs.values())) == 1, f"modes disagree on nnz: {nnzs}"
    fam_names = cb._capi.KERNEL_FAMILIES
    out = {"config": config, "validated_rows": validated, "modes": {}}
    for m, v in results.items():
        agg = v["agg"]
        out["modes"][m] = {
            "ms_per_step": v["ms_per_step"], "products_per_s": v["products_per_s"], "nnz_per_s": v["nnz_per_s"],
            "row_blocks": agg["blocks"], "replans": agg["replans"], "plan_ms": agg["plan_ms"],
            "ms_kernel": dict(zip(fam_names, agg["ms_kernel"])), "launches": dict(zip(fam_names, agg["launches"])),
            "fallback_rows": agg["fallback_rows"], "fallback_rate": agg["fallback_rows"] / max(1, int(rows.size)),
            "pred_abs_log2_err": agg["err_sum"] / agg["err_rows"] if m == "learned" and agg["err_rows"] else None,
            "capacity_over_nnz": agg["capacity"] / max(1, agg["nnz"]) if m != "plain" else None,
            "dense_rows": agg["dense_rows"], "dense_nnz": agg["dense_nnz"], "tuning": tuned[m],
            "pool_in_timed_steps": agg["pool_timed"],
            # per bin (0 empty, 1 tiny, 2-9 hash 2^7..2^14 slots, 10-12 LDS ESC,
            # 13-31 bitmap by product count), last timed step
            "bin_rows": agg["bin_rows"], "bin_products": agg["bin_products"],
            "roofline": roofline_of(cb, agg, workload),
            "step_hbm_frac": step_algorithmic_bytes(int(rows.size), int(cols.size), total_products,
                                                    int(rows.size), agg["nnz"]) / (v["ms_per_step"] / 1e3) / 1e9
            / HBM_PEAK_GBS}
    out["nnz_out"] = int(next(iter(nnzs.values())))
    for x, y in (("learned", "plain"), ("bound", "plain"), ("learned", "bound")):
        if x in results and y in results:
            out[f"{x}_vs_{y}"] = results[x]["products_per_s"] / results[y]["products_per_s"]
    return out, A


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22", choices=sorted(WORKLOADS))
    ap.add_argument("--also", default="rmat20", help="comma-separated workloads reported beside the main one "
                    "('' for none)")
    ap.add_argument("--mode", default="all", choices=["all", "plain", "learned", "bound"])
    ap.add_argument("--block-bound", type=float, default=float((1 << 32) - 1),
                    help="largest result of one row block (nonzeros)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--closure-iters", type=int, default=4,
                    help="steps of the 64-label closure workload reported beside the main metric (0: skip)")
    ap.add_argument("--tuning-file", default=None,
                    help="JSON of knobs per mode (default profiles/tuning_<workload>.json when it exists)")
    ap.add_argument("--set", action="append", default=[], metavar="KNOB=VALUE",
                    help="override a tuning knob for every mode (after the tuning file)")
    ap.add_argument("--autotune", action="store_true", help="tune the knobs on the workload first")
    ap.add_argument("--save-tuning", default=None)
    ap.add_argument("--hypersparse", type=int, default=200,
                    help="also run the hypersparse learned-sizing workload, training on this many instances")
    ap.add_argument("--hypersparse-only", type=int, default=0,
                    help="run only the hypersparse workload, training on this many instances")
    ap.add_argument("--closure-only", type=int, default=0,
                    help="run only the 64-label closure, up to this many steps or its fixed point, per-triple "
                         "plain against batched learned (each after a warm-up run of its own)")
    ap.add_argument("--closure-cold", action="store_true", help="with --closure-only: no warm-up runs")
    ap.add_argument("--closure-legs", default="per_triple_plain,batched_learned",
                    help="with --closure-only: comma-separated legs (per_triple_plain, batched_plain, batched_learned)")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the automaton (x) R-MAT 22/32 workload reported beside the main metric")
    ap.add_argument("--c5-only", action="store_true",
                    help="run only the automaton (x) R-MAT 22/32 workload: the Kronecker in automaton-row "
                         "blocks and a bounded sample of one K * K closure step")
    ap.add_argument("--copy-probe", type=float, default=0,
                    help="only run a device-to-device copy of this many bytes (4 times): the reference "
                         "point that scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.copy_probe:
        import clbool_amd as cb
        ctl = cb.create_controls()
        nbytes = int(args.copy_probe)
        src = cb.Buffer.create(ctl, nbytes)
        dst = cb.Buffer.create(ctl, nbytes)
        cb.fill_with(ctl, src, nbytes // 4, 1)
        for _ in range(4):
            t0 = time.perf_counter()
            cb.lib().clb_buffer_copy(ctl.handle, src.handle, 0, dst.handle, 0, nbytes)
            dt = time.perf_counter() - t0
            log(f"[copy-probe] {nbytes} bytes in {dt * 1e3:.3f} ms ({2 * nbytes / dt / 1e9:.1f} GB/s read+write)")
        print(json.dumps({"copy_probe_bytes": nbytes}))
        return
    if args.c5_only:
        import clbool_amd as cb
        from clbool_amd import generators as gen
        ctl = cb.create_controls()
        print(json.dumps({"c5": bench_c5(cb, gen, ctl, log)}), flush=True)
        return
    if args.closure_only:
        import clbool_amd as cb
        from clbool_amd import generators as gen
        ctl = cb.create_controls()
        train_predictor(cb, gen, ctl, log)
        print(json.dumps({"closure": bench_closure(cb, gen, ctl, args.closure_only, log,
                                                   legs=tuple(args.closure_legs.split(",")),
                                                   warm=not args.closure_cold)}),
              flush=True)
        return
    if args.hypersparse_only:
        import clbool_amd as cb
        from clbool_amd import generators as gen
        ctl = cb.create_controls()
        print(json.dumps({"hypersparse": bench_hypersparse(cb, gen, ctl, args.hypersparse_only, log)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU row-block sharding is a "
              "design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import clbool_amd as cb
    from clbool_amd import autotune as at
    from clbool_amd import generators as gen
    import oracle  # checker only: validates sampled rows and times the CPU baseline

    ctl = cb.create_controls()
    dev = cb.show_devices()[0]
    modes = ["plain", "learned", "bound"] if args.mode == "all" else [args.mode]
    train = train_predictor(cb, gen, ctl, log) if "learned" in modes else None
    base_tuning = at.tuning_dict(ctl.tuning())
    tpath = args.tuning_file or os.path.join(ROOT, "profiles", f"tuning_{args.workload}.json")
    loaded = json.load(open(tpath)) if os.path.exists(tpath) else {}
    overrides = {k: int(v) for k, v in (kv.split("=", 1) for kv in args.set)}
    for k in overrides:
        if k not in at.KNOBS:
            raise SystemExit(f"unknown knob {k}; knobs: {sorted(at.KNOBS)}")
    tuned = {m: {**base_tuning, **loaded.get(m, {}), **overrides} for m in modes}
    if loaded:
        log(f"[bench] knobs from {os.path.relpath(tpath, ROOT)}")
    block_bound = int(args.block_bound)

    if args.autotune:
        scale, ef, seed = WORKLOADS[args.workload]
        r, c = gen.rmat(scale, ef, seed)
        rpt, rows, cols = gen.to_dcsr(r, c)
        a_t = cb.matrix_dcsr_from_cpu(ctl, rpt, rows, cols, 1 << scale, 1 << scale)
        for m in modes:
            at.apply(ctl, tuned[m])
            blocks = plan_blocks(cb, ctl, a_t, MODE_CODES[m], block_bound)
            res = at.autotune(ctl, a_t, a_t, blocks, mode=MODE_CODES[m], repeats=2, log=log)
            tuned[m] = res["tuning"]
            log(f"[bench] {m}: autotuned {res['tuning']} over {res['evaluated']} settings")
        del a_t
        if args.save_tuning:
            with open(args.save_tuning, "w") as f:
                json.dump(tuned, f)

    main_res, A = bench_product(cb, gen, ctl, oracle, args.workload, modes, args.steps, args.warmup, tuned,
                                not args.no_validate, block_bound, log)
    best = "learned" if "learned" in main_res["modes"] else modes[0]
    R = main_res["modes"][best]
    out = {
        "metric": METRIC, "value": R["products_per_s"], "unit": "products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {**main_res["config"], "nnz_out": main_res["nnz_out"], "row_blocks": R["row_blocks"],
                   "mode": best, "device": dev["arch"], "compute_units": dev["compute_units"]},
        "output_nnz_per_s": R["nnz_per_s"],
        "roofline": R["roofline"],
        "predictor_training": train,
        "pred_abs_log2_err": R.get("pred_abs_log2_err"),
        "fallback_rate": R.get("fallback_rate"),
    }
    for k in ("learned_vs_plain", "bound_vs_plain", "learned_vs_bound", "modes", "validated_rows"):
        if k in main_res:
            out[k] = main_res[k]
    for w in [w for w in args.also.split(",") if w and w != args.workload]:
        wt = os.path.join(ROOT, "profiles", f"tuning_{w}.json")
        wl = json.load(open(wt)) if os.path.exists(wt) else {}
        tw = {m: {**base_tuning, **wl.get(m, {}), **overrides} for m in modes}
        res, _ = bench_product(cb, gen, ctl, oracle, w, modes, args.steps, args.warmup, tw, not args.no_validate,
                               block_bound, log)
        out[w] = res
    if args.closure_iters > 0 and "learned" in modes:
        at.apply(ctl, tuned.get("learned", base_tuning))
        out["closure"] = bench_closure(cb, gen, ctl, args.closure_iters, log)
    if args.hypersparse > 0:
        out["hypersparse"] = bench_hypersparse(cb, gen, ctl, args.hypersparse, log)
    if not args.no_c5 and "learned" in modes:
        ctl.trim()
        # each mode's step with that mode's knobs of the main workload (K is
        # 2^22 columns wide, as R-MAT 22)
        out["c5"] = bench_c5(cb, gen, ctl, log,
                             tunings={m: tuned.get(m, base_tuning) for m in ("plain", "learned")})
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(oracle, A)
        try:
            nproc = len(os.sched_getaffinity(0))
        except AttributeError:
            nproc = os.cpu_count() or 1
        threads = int(os.environ.get("OMP_NUM_THREADS", nproc))
        out["cpu_gustavson"] = cpu_gustavson(oracle, A, threads)
        out["host"] = {"nproc": nproc, "os_cpu_count": os.cpu_count(), "omp_num_threads": threads}
        out["clbool_cpu_opencl"] = ("not run: clBool's create_controls asks only for GPU devices "
                                    "(src/common/env.cpp:21) and no CPU OpenCL runtime is installed")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
