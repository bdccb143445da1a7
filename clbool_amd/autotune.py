"""Autotuner for the SpGEMM's bin thresholds, table sizes and kernel variants.

Replaces the hard-coded thresholds of clBool's hash SpGEMM
(src/dcsr/dcsr_matrix_multiplication_hash.cpp:13-47).  The tuner times the
library on a sample of the caller's own workload (row blocks of A * B) and
walks the knobs one at a time (coordinate descent).  A setting is kept only
if every block's result is the same as with the starting knobs (sizes and the
device-side digest of the whole matrix, clb_dcsr_digest) and its best time
over the repeats beats the current best by at least `min_gain` (2% by
default: smaller differences are within run-to-run noise).
The winning knobs are stored per device (architecture and CU count) in a JSON
file and can be reloaded without tuning again.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

KNOBS = {
    "emit_lane_bits_min": [256, 1024, 4096],   # 4096: never one lane per bit
    "dump_bits_per_chunk": [64, 256, 1024, 4096, 65536],   # 65536: no dump
    "bitmap_persistent": [0, 1],
    "bitmap_min_products": [4096, 8192, 16384, 32768, 65536],
    "hash_max_slots": [4096, 8192, 16384],
    "bitmap_slot_ratio": [1, 2, 4],
    "tiny_max_products": [16, 32, 64],
    "learned_margin_pct": [10, 25, 50],
    "bitmap_pred_caps": [0, 1],   # one pass: bitmap rows sized by the prediction (1) or the bound (0)
    "gather_combine": [0, 1],     # bitmap gather: one LDS atomic per word and wave (1) or per product (0)
    "esc_max_products": [0, 2048, 8192],  # multi-window bitmap rows up to this many products sort in LDS
    "compact_tiles": [0, 1],      # one pass: compaction by output tiles (1) or one workgroup per row (0)
    "onepass_bitmap_min": [0, 8192, 32768, 131072],  # one pass: product count that forces a bitmap (0: never)
    "bitmap_window_bits": [0, 13, 14, 15],  # numeric / one-pass bitmap window, 2^bits words (0: by width)
    "emit_sparse_bits": [0, 64, 1024, 4096],  # bitmap slices below this many bits per 1024 words: lane-own emission
}
# workgroup shape of each LDS hash table size: default (0) or twice the threads (1)
for _b in range(7, 14):
    KNOBS[f"hash_wide_{_b}"] = [0, 1]
# knobs that live in clb_tuning.reserved[] (whole words, or one bit of a word)
RESERVED = {"bitmap_pred_caps": 0, "gather_combine": 1, "esc_max_products": 2, "compact_tiles": 3,
            "onepass_bitmap_min": 5, "bitmap_window_bits": 6, "emit_sparse_bits": 7}
RESERVED_BITS = {f"hash_wide_{b}": (4, b - 7) for b in range(7, 14)}
# on matrices wider than one bitmap window: rows within the largest hash table
# take a table rather than a bitmap of several column slices (1)
KNOBS["tables_over_slices"] = [0, 1]
RESERVED_BITS["tables_over_slices"] = (4, 16)
LEARNED_ONLY = ("learned_margin_pct", "bitmap_pred_caps")   # knobs of the predictor's sizing
ONE_PASS_ONLY = ("compact_tiles", "onepass_bitmap_min")      # knobs of the learned and bound one pass


def _get(t, k):
    if k in RESERVED_BITS:
        w, b = RESERVED_BITS[k]
        return (t.reserved[w] >> b) & 1
    return t.reserved[RESERVED[k]] if k in RESERVED else getattr(t, k)


def _set(t, k, v):
    if k in RESERVED_BITS:
        w, b = RESERVED_BITS[k]
        t.reserved[w] = (t.reserved[w] & ~(1 << b)) | ((int(v) & 1) << b)
    elif k in RESERVED:
        t.reserved[RESERVED[k]] = v
    else:
        setattr(t, k, v)


def tuning_dict(t) -> dict:
    return {k: int(_get(t, k)) for k in KNOBS}


def apply(controls, d: dict):
    t = controls.tuning()
    for k, v in d.items():
        if k in KNOBS:
            _set(t, k, int(v))
    controls.set_tuning(t)


def _checksum(cb, controls, c) -> tuple:
    """(nnz, nzr, digest of the whole result): a setting that changes any
    entry of any row changes the digest."""
    if c.nnz == 0:
        return (0, 0, 0)
    return (int(c.nnz), int(c.nzr), cb.dcsr.digest(controls, c))


def accept(t_new, t_cur, sums_new, sums_ref, min_gain=0.02, what=""):
    """Decision of one trial: raises if the result differs from the
    reference; True if the trial is at least min_gain faster."""
    if sums_new != sums_ref:
        raise AssertionError(f"tuning {what} changed the result")
    return t_new < (1.0 - min_gain) * t_cur


def autotune(controls, a, b, blocks, mode=0, repeats=3, knobs=None, log=None, passes=2, min_gain=0.02) -> dict:
    """Coordinate descent over KNOBS on the given row blocks of A * B (the
    knobs interact, so the sweep over all of them runs `passes` times)."""
    import clbool_amd as cb

    def relevant(k):
        if k in LEARNED_ONLY:
            return mode == cb.SPGEMM_LEARNED
        if k in ONE_PASS_ONLY:
            return mode in (cb.SPGEMM_LEARNED, cb.SPGEMM_BOUND)
        return True

    knobs = knobs or [k for k in KNOBS if relevant(k)]

    def run():
        best = None
        sums = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            sums = []
            for lo, hi in blocks:
                c, _ = cb.dcsr.spgemm(controls, a, b, mode=mode, row_begin=lo, row_end=hi)
                sums.append(_checksum(cb, controls, c))
                del c
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best, sums

    cur = tuning_dict(controls.tuning())
    run()  # warm the caches and the allocator
    t_cur, ref = run()
    history = [(dict(cur), t_cur)]
    for k in [k for _ in range(passes) for k in knobs]:
        for v in KNOBS[k]:
            if v == cur[k]:
                continue
            trial = dict(cur)
            trial[k] = v
            try:
                apply(controls, trial)
            except Exception:
                continue
            t, sums = run()
            history.append((dict(trial), t))
            if log:
                log(f"[autotune] {k}={v}: {t * 1e3:.2f} ms (best {t_cur * 1e3:.2f})")
            if accept(t, t_cur, sums, ref, min_gain, str(trial)):
                cur, t_cur = trial, t
        apply(controls, cur)
    return {"tuning": cur, "seconds": t_cur, "evaluated": len(history)}


def cache_path(arch: str, cus: int, mode: str) -> str:
    root = os.environ.get("CLBOOL_TUNING_DIR", os.path.join(os.path.expanduser("~"), ".cache", "clbool_amd"))
    return os.path.join(root, f"tuning_{arch.split(':')[0]}_{cus}cu_{mode}.json")


def save(path: str, result: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(controls, path: str) -> bool:
    if not os.path.exists(path):
        return False
    with open(path) as f:
        apply(controls, json.load(f)["tuning"])
    return True
