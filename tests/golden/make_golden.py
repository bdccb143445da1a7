"""Writes the golden fixtures of tests/golden/.

* fastrandom.json -- draws of the reference's own FastRandom
  (libs/utils/libutils/fast_random.h, compiled from the clBool tree into
  oracle/_ref by oracle/Makefile); the oracle's restatement must match them.
* known_answers.json -- hand-derived answers of the reference's own test
  inputs: the README example product (test/test_operations.cpp:284-291), the
  COO constructor dedup case (test/test_procedures.cpp:20-31) and the
  csr_cpu_from_pairs vector (test/check_test_utils.cpp:21-39).
* generated_pairs.json -- size and checksums of generate_coo_pairs_cpu
  (src/coo/coo_utils.cpp:204-218) outputs for the reference test sizes,
  computed by the oracle's restatement (coo_utils.cpp itself does not build
  standalone: its includes reach the CMake-generated headers_map.hpp).

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref with the clBool tree present)")
    fr = {"source": "libs/utils/libutils/fast_random.h via oracle/_ref", "cases": []}
    for seed, lo, hi in [(10000, 0, 1023), (150000, 0, 2999), (190000, 0, 3799), (0, 0, 9), (1, -5, 5),
                         (123456789, 0, 2147483647)]:
        d = oracle.ref_fastrandom_stream(seed, lo, hi, 64)
        fr["cases"].append({"seed": seed, "min": lo, "max": hi, "draws": [int(x) for x in d]})
    with open(os.path.join(HERE, "fastrandom.json"), "w") as f:
        json.dump(fr, f, indent=1)

    ka = {
        "readme_product": {
            "source": "test/test_operations.cpp:284-291 (inputs); answer derived by hand",
            "a": {"rows": [0, 0, 0, 2, 2, 4], "cols": [0, 1, 4, 2, 3, 2], "nrows": 5, "ncols": 5},
            "b": {"rows": [1, 1, 2, 3, 3, 3, 5], "cols": [0, 4, 2, 2, 3, 4, 2], "nrows": 5, "ncols": 5},
            "c": {"rpt": [0, 2, 5, 6], "rows": [0, 2, 4], "cols": [0, 4, 2, 3, 4, 2]},
        },
        "coo_constructor": {
            "source": "test/test_procedures.cpp:20-31",
            "rows": [0, 0, 1, 1, 3, 4, 4, 5, 5, 5, 6], "cols": [1, 2, 0, 3, 5, 2, 2, 0, 3, 3, 5],
            "nrows": 7, "ncols": 7, "nnz_after": 9,
            "pairs_after": [[0, 1], [0, 2], [1, 0], [1, 3], [3, 5], [4, 2], [5, 0], [5, 3], [6, 5]],
        },
        "csr_cpu_from_pairs": {
            "source": "test/check_test_utils.cpp:21-39",
            "pairs": [[0, 0], [0, 1], [0, 3], [2, 1], [2, 3], [2, 5], [3, 1], [3, 4], [3, 5]],
            "m": 5, "n": 7, "rpt": [0, 3, 3, 6, 9, 9], "cols": [0, 1, 3, 1, 3, 5, 1, 4, 5],
        },
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(ka, f, indent=1)

    gp = {"source": "oracle restatement of generate_coo_pairs_cpu (coo_utils.cpp:204-218)", "cases": []}
    for size, k in [(1024, 10), (3000, 50), (3800, 55), (200, 3), (15, 7)]:
        r, c = oracle.generate_coo_pairs(size * k, size)
        key = (r.astype(np.uint64) << np.uint64(32)) | c.astype(np.uint64)
        gp["cases"].append({"size": size, "k": k, "n": int(r.size),
                            "sum_rows": int(r.astype(np.uint64).sum()), "sum_cols": int(c.astype(np.uint64).sum()),
                            "xor_keys": int(np.bitwise_xor.reduce(key)) if key.size else 0,
                            "first": [[int(r[i]), int(c[i])] for i in range(min(8, r.size))]})
    with open(os.path.join(HERE, "generated_pairs.json"), "w") as f:
        json.dump(gp, f, indent=1)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
