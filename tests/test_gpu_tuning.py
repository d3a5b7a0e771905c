"""Launch-schedule autotuner: every schedule (lanes per row of the grouped
sweeps, entries per lane of the compaction tiles, rows per workgroup of the
row folds) presolves to bit-identical results, the tuner returns a valid
schedule, and invalid schedules are rejected."""
import numpy as np
import pytest

import gpupre as G

pytestmark = pytest.mark.gpu


def lp_zipf():
    w = dict(G.WORKLOADS["medium"])
    w.pop("m"), w.pop("n")
    return G.generate(40_000, 60_000, 3, **w)


def run(d):
    s = G.preprocess(d, G.parity_params())
    try:
        return s.record(), s.reduced(), s.obj_constant_delta
    finally:
        s.free()


def same(a, b):
    (ra, la, oa), (rb, lb, ob) = a, b
    assert np.array_equal(ra.row_red2org, rb.row_red2org) and np.array_equal(ra.col_red2org, rb.col_red2org)
    assert np.array_equal(ra.types, rb.types) and np.array_equal(ra.indices, rb.indices)
    assert np.array_equal(ra.vals.view(np.uint64), rb.vals.view(np.uint64))
    assert np.array_equal(la.vals.view(np.uint64), lb.vals.view(np.uint64)) and oa == ob


@pytest.mark.parametrize("lanes,per_lane,rows", [(2, 4, 128), (8, 16, 512), (16, 8, 256), (64, 4, 128)])
def test_schedules_give_identical_results(lanes, per_lane, rows):
    lp = lp_zipf()
    with G.DeviceLP(lp) as d:
        G.set_tuning()
        ref = run(d)
        G.set_tuning(G.GpTuning(lanes, lanes, per_lane, lanes, rows))
        try:
            got = run(d)
        finally:
            G.set_tuning()
    same(ref, got)


def test_autotune_returns_valid_schedule_and_keeps_current():
    lp = lp_zipf()
    with G.DeviceLP(lp) as d:
        G.set_tuning()
        best, tuned, default = G.autotune(d.lp, G.parity_params(), reps=1)
        cur = G.GpTuning()
        G.check(G.load().gp_get_tuning(cur))
        assert cur.as_dict() == G.GpTuning().as_dict()
        for k in ("lanes_candidates", "lanes_col_max", "lanes_live_stats"):
            assert getattr(best, k) in (0, 2, 4, 8, 16, 64)
        assert best.compact_per_lane in (0, 4, 8, 16)
        assert best.fold_rows in (0, 128, 256, 512)
        assert 0 < tuned <= default
        G.set_tuning(best)
        try:
            got = run(d)
        finally:
            G.set_tuning()
        same(run(d), got)


def test_invalid_schedule_rejected():
    with pytest.raises(G.GpError):
        G.set_tuning(G.GpTuning(3, 0, 0, 0, 0))
    with pytest.raises(G.GpError):
        G.set_tuning(G.GpTuning(0, 0, 0, 0, 300))
    with pytest.raises(G.GpError):
        G.set_tuning(G.GpTuning(0, 0, 2, 0, 0))
