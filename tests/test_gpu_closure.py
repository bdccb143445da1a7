"""Closure steps X |= Y * Z over label matrices (the CFPQ/RPQ workload):
the per-triple and the batched (stacked-operand) modes, in plain and learned
SpGEMM modes, against the oracle's SpGEMM and union step by step."""
import numpy as np
import pytest

import clbool_amd as cb
import oracle
from clbool_amd import generators as gen
from helpers import assert_dcsr_equal

pytestmark = pytest.mark.gpu


def _oracle_step(mats, triples):
    """The step on the CPU: products from the start-of-step matrices, then unions."""
    prods = [oracle.spgemm(mats[y], mats[z]) for _, y, z in triples]
    new = list(mats)
    for (x, _, _), p in zip(triples, prods):
        X = new[x]
        xr = np.repeat(X.rows, np.diff(X.rpt.astype(np.int64)))
        pr = np.repeat(p.rows, np.diff(p.rpt.astype(np.int64)))
        r, c = oracle.coo_add(xr, X.cols, pr, p.cols)
        new[x] = oracle.pairs_to_dcsr(r, c, X.nrows, X.ncols)
    return new


def _labels(n_labels, log_n, lo, hi, seed):
    n = 1 << log_n
    out = []
    for r, c in gen.closure_labels(n_labels=n_labels, log_n=log_n, lo=lo, hi=hi, seed=seed):
        out.append(oracle.pairs_to_dcsr(r, c, n, n))
    return out, n


@pytest.mark.parametrize("mode,block_bound", [(cb.CLOSURE_PER_TRIPLE, 0), (cb.CLOSURE_BATCHED, 0),
                                              (cb.CLOSURE_BATCHED, 20000)])
@pytest.mark.parametrize("spgemm_mode", [cb.SPGEMM_PLAIN, cb.SPGEMM_LEARNED])
def test_closure_steps_match_oracle(controls, mode, block_bound, spgemm_mode):
    cpu, n = _labels(12, 10, 1 << 6, 1 << 11, seed=4)
    triples = gen.closure_triples(12, 8, seed=5) + [(0, 1, 2), (0, 3, 4)]  # X repeated
    mats = [cb.matrix_dcsr_from_cpu(controls, m.rpt, m.rows, m.cols, n) for m in cpu]
    if spgemm_mode == cb.SPGEMM_LEARNED:
        cb.predictor_reset(controls, 9)
    for step in range(3):
        st = cb.closure_step(controls, mats, triples, mode=mode, spgemm_mode=spgemm_mode, block_bound=block_bound)
        cpu = _oracle_step(cpu, triples)
        for k in range(len(cpu)):
            assert_dcsr_equal(controls, mats[k], cpu[k], f"step {step} matrix {k}")
        if mode == cb.CLOSURE_PER_TRIPLE:
            assert st["spgemm_calls"] == len(triples)
        elif block_bound == 0:
            assert st["spgemm_calls"] == 1
        else:  # the stacked product is cut into several row blocks, each
            # within the bound (rows straddle multiples of the cut spacing)
            assert st["spgemm_calls"] > 1
            assert 0 < st["max_block_bound"] <= block_bound


def test_closure_reaches_fixed_point(controls):
    cpu, n = _labels(4, 8, 1 << 4, 1 << 7, seed=6)
    triples = [(0, 0, 1), (1, 1, 2), (2, 2, 3), (3, 3, 0)]
    mats = [cb.matrix_dcsr_from_cpu(controls, m.rpt, m.rows, m.cols, n) for m in cpu]
    for it in range(60):
        st = cb.closure_step(controls, mats, triples, mode=cb.CLOSURE_BATCHED)
        cpu = _oracle_step(cpu, triples)
        if st["changed"] == 0:
            break
    assert st["changed"] == 0, "no fixed point within 60 steps"
    for k in range(4):
        assert_dcsr_equal(controls, mats[k], cpu[k], f"fixed point matrix {k}")
    # one more step changes nothing
    assert cb.closure_step(controls, mats, triples, mode=cb.CLOSURE_PER_TRIPLE)["changed"] == 0
