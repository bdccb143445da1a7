"""Host-side logic of the autotuner: the accept rule and the knob table.
No GPU calls."""
import pytest

from clbool_amd import autotune as at


def test_accept_requires_identical_results():
    ref = [(10, 3, 123), (7, 2, 55)]
    assert at.accept(0.5, 1.0, list(ref), ref)
    # a middle block whose digest differs is a wrong result, however fast
    with pytest.raises(AssertionError):
        at.accept(0.5, 1.0, [(10, 3, 123), (7, 2, 56)], ref)
    with pytest.raises(AssertionError):
        at.accept(0.5, 1.0, [(10, 3, 123)], ref)


def test_accept_requires_the_minimum_gain():
    ref = [(1, 1, 1)]
    assert not at.accept(0.99, 1.0, ref, ref)          # 1% faster: noise
    assert not at.accept(0.98, 1.0, ref, ref)          # exactly 2%: not more
    assert at.accept(0.979, 1.0, ref, ref)
    assert at.accept(0.99, 1.0, ref, ref, min_gain=0.005)
    assert not at.accept(1.2, 1.0, ref, ref)


class _T:
    def __init__(self):
        self.reserved = [0] * 8
        for k in at.KNOBS:
            if k not in at.RESERVED:
                setattr(self, k, 0)


def test_knobs_round_trip_including_reserved_slots():
    t = _T()
    for k, vals in at.KNOBS.items():
        at._set(t, k, vals[-1])
    d = {k: at._get(t, k) for k in at.KNOBS}
    assert d == {k: v[-1] for k, v in at.KNOBS.items()}
    # the workgroup-shape bits and the tables-over-slices bit share one word
    # without disturbing each other
    assert t.reserved[4] == 0x1007F
    at._set(t, "hash_wide_9", 0)
    assert t.reserved[4] == 0x1007B and at._get(t, "hash_wide_10") == 1
    at._set(t, "tables_over_slices", 0)
    assert t.reserved[4] == 0x7B and at._get(t, "hash_wide_13") == 1
    assert t.reserved[at.RESERVED["bitmap_pred_caps"]] == at.KNOBS["bitmap_pred_caps"][-1]
    assert t.reserved[at.RESERVED["gather_combine"]] == at.KNOBS["gather_combine"][-1]
