"""Synthetic workload generators (host, numpy) for the SpGEMM configurations.

These build the inputs of the benchmark and test workloads; they are not
part of the GPU path.  All take an explicit seed.

* ``uniform_pairs``: n distinct (row, col) pairs drawn uniformly (the
  1024 x 1024 / 10,000-pair configuration).
* ``rmat``: Graph500-style R-MAT edges (a, b, c, d) = (.57, .19, .19, .05),
  duplicates and self-loops removed.
* ``hypersparse_zipf``: 2^24 x 2^24 with 1% nonempty rows, Zipf(1.5) row
  lengths capped at 4096, uniform columns.
* ``closure_labels``: label matrices on one vertex set with Zipf(1.1) sizes.
* ``to_dcsr``: sorted unique pairs -> (rpt, rows, cols).
"""
from __future__ import annotations

import numpy as np


def _unique_sorted(rows, cols):
    key = (rows.astype(np.uint64) << np.uint64(32)) | cols.astype(np.uint64)
    key = np.unique(key)
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def to_dcsr(rows, cols):
    """Sorted unique (row, col) pairs -> DCSR arrays (rpt, rows, cols)."""
    rows = np.asarray(rows, dtype=np.uint32)
    cols = np.asarray(cols, dtype=np.uint32)
    if rows.size == 0:
        return np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    starts = np.flatnonzero(np.r_[True, rows[1:] != rows[:-1]])
    rpt = np.r_[starts, rows.size].astype(np.uint32)
    return rpt, rows[starts].astype(np.uint32), cols.astype(np.uint32)


def uniform_pairs(n, nrows, ncols, seed):
    """n distinct pairs uniform over nrows x ncols (draw until n are distinct)."""
    if n > nrows * ncols:
        raise ValueError("more pairs than cells")
    rng = np.random.default_rng(seed)
    have = np.zeros(0, np.uint64)
    while have.size < n:
        m = int((n - have.size) * 1.2) + 16
        r = rng.integers(0, nrows, m, dtype=np.uint64)
        c = rng.integers(0, ncols, m, dtype=np.uint64)
        have = np.unique(np.r_[have, (r << np.uint64(32)) | c])
    if have.size > n:  # keep a uniform subset of exactly n
        have = np.sort(rng.choice(have, n, replace=False))
    return (have >> np.uint64(32)).astype(np.uint32), (have & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def _rmat_bits(x, a, ab, abc):
    """Row and column ids from the per-level uniform draws x[level, edge]."""
    k = x.shape[1]
    r = np.zeros(k, np.uint32)
    q = np.zeros(k, np.uint32)
    for lv in range(x.shape[0]):
        xl = x[lv]
        rb = xl >= ab
        cb = np.where(rb, xl >= abc, xl >= a)
        r = (r << np.uint32(1)) | rb.astype(np.uint32)
        q = (q << np.uint32(1)) | cb.astype(np.uint32)
    keep = r != q
    return (r[keep].astype(np.uint64) << np.uint64(32)) | q[keep].astype(np.uint64)


def _rmat_chunk(state, skip, scale, k, a, ab, abc):
    """The chunk whose draws start `skip` 64-bit outputs into the stream."""
    bg = np.random.PCG64()
    bg.state = state
    bg.advance(skip)
    return _rmat_bits(np.random.Generator(bg).random((scale, k), dtype=np.float32), a, ab, abc)


def rmat(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, chunk=1 << 22, n_edges=None, threads=8):
    """R-MAT graph on 2^scale vertices: sorted unique pairs without self-loops.
    Edge e of the chunk starting at edge s0 (chunk length k) takes level l's
    quadrant from float32 draw number s0 * scale + l * k + (e - s0) of the
    seeded PCG64 stream.  A float32 draw takes half of one 64-bit output, and
    every chunk before the last holds an even number of draws, so each chunk
    starts its own copy of the stream advanced to its first draw and the
    chunks are drawn in parallel with the same result as drawing in order."""
    import concurrent.futures as cf
    rng = np.random.default_rng(seed)
    m = int(n_edges if n_edges is not None else edge_factor * (1 << scale))
    ab, abc = a + b, a + b + c
    starts = list(range(0, m, chunk))
    state = rng.bit_generator.state
    if (scale * chunk) % 2:
        raise ValueError("chunks must hold an even number of draws")
    jobs = [(state, s0 * scale // 2, scale, min(chunk, m - s0), a, ab, abc) for s0 in starts]
    if len(jobs) <= 1:
        keys = [_rmat_chunk(*j) for j in jobs]
    else:
        with cf.ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
            keys = list(ex.map(lambda j: _rmat_chunk(*j), jobs))
    key = np.unique(np.concatenate(keys)) if keys else np.zeros(0, np.uint64)
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def hypersparse_zipf(seed, log_n=24, row_frac=0.01, s=1.5, cap=4096):
    """2^log_n square matrix, row_frac of rows nonempty, Zipf(s) lengths <= cap."""
    rng = np.random.default_rng(seed)
    n = 1 << log_n
    nzr = int(n * row_frac)
    rows = np.sort(rng.choice(n, nzr, replace=False)).astype(np.uint32)
    lens = np.minimum(rng.zipf(s, nzr), cap).astype(np.int64)
    total = int(lens.sum())
    cols = rng.integers(0, n, total, dtype=np.uint64)
    rr = np.repeat(rows.astype(np.uint64), lens)
    key = np.unique((rr << np.uint64(32)) | cols)
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def bounded_zipf_sizes(k, lo, hi, s, seed):
    """k sizes in [lo, hi] with density proportional to n^-s (inverse CDF)."""
    rng = np.random.default_rng(seed)
    u = rng.random(k)
    e = 1.0 - s
    lo_e, hi_e = float(lo) ** e, float(hi) ** e
    return np.floor((lo_e + u * (hi_e - lo_e)) ** (1.0 / e)).astype(np.int64).clip(lo, hi)


def closure_labels(n_labels=64, log_n=18, lo=1 << 10, hi=1 << 20, s=1.1, seed=0):
    """Label matrices (pairs) on one vertex set of 2^log_n with Zipf(s) sizes."""
    sizes = bounded_zipf_sizes(n_labels, lo, hi, s, seed)
    mats = []
    for i, m in enumerate(sizes):
        mats.append(rmat(log_n, 0, seed * 1000 + i + 1, n_edges=int(m)))
    return mats


def closure_triples(n_labels, n_triples, seed):
    """n_triples (X, Y, Z) label indices for the X += Y * Z closure step."""
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.integers(0, n_labels, 3)) for _ in range(n_triples)]
