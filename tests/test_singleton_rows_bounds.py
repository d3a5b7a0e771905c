"""Singleton rows: the bound check covers every column, not only the columns
of singleton rows.

The reference copies the plan's bounds into candidate arrays, tightens them
from the live singleton rows and then checks every column j < n for crossed
candidates and for a change (rule_singleton_rows.cu:104-127, 164-214).  A
column no singleton row touches therefore still reports crossed bounds as
infeasible, and a NaN bound as a change.  The engine seeds candidates only
for the columns of singleton rows and reads the current bounds of every other
column in its check, so these cases pin that the two agree.
"""
import numpy as np
import pytest

import gpupre as G
from backends import make_backend

INF = np.inf


def lp_case(l, u, AL, AU):
    # rows 0 and 1 are singleton rows on column 0 and column 1; row 2 has three entries
    dense = np.array([[2.0, 0.0, 0.0, 0.0],
                      [0.0, -1.0, 0.0, 0.0],
                      [1.0, 1.0, 1.0, 1.0]])
    m, n = dense.shape
    return G.HostLP.from_dense(dense, np.zeros(n), np.asarray(l, float), np.asarray(u, float),
                               np.asarray(AL, float), np.asarray(AU, float))


CASES = {
    # tightening on columns 0 and 1; columns 2 and 3 untouched and consistent
    "tighten": dict(l=[-10, -10, 0, 0], u=[10, 10, 5, 5], AL=[-4, -3, -INF], AU=[4, 6, 8]),
    # column 3, touched by no singleton row, has crossed bounds: infeasible
    "crossed_untouched": dict(l=[-10, -10, 0, 3], u=[10, 10, 5, 1], AL=[-4, -3, -INF], AU=[4, 6, 8]),
    # column 2, untouched, has a NaN upper bound: the check reports a change
    "nan_untouched": dict(l=[-10, -10, 0, 0], u=[10, 10, np.nan, 5], AL=[-20, -30, -INF], AU=[20, 30, 8]),
    # the singleton rows imply crossed bounds on column 0
    "crossed_implied": dict(l=[-10, -10, 0, 0], u=[10, 10, 5, 5], AL=[4, -3, -INF], AU=[2, 6, 8]),
}


def run(kind, case):
    b = make_backend(kind, lp_case(**case))
    try:
        b.set_live(0)
        b.refresh()
        b.apply("singleton_rows", G.default_params())
        out = {k: b.get(k) for k in ("keep_row_mask", "new_l", "new_u")}
        out.update(b.flags())
        return out
    finally:
        if hasattr(b, "free"):
            b.free()


def test_restatement_checks_every_column():
    crossed = run("oracle", CASES["crossed_untouched"])
    assert crossed["has_infeasible"] == 1
    nan = run("oracle", CASES["nan_untouched"])
    assert nan["has_infeasible"] == 0 and nan["has_change"] == 1
    ok = run("oracle", CASES["tighten"])
    assert ok["has_infeasible"] == 0
    assert list(ok["new_l"][:2]) == [-2.0, -6.0] and list(ok["new_u"][:2]) == [2.0, 3.0]
    assert list(ok["keep_row_mask"]) == [0, 0, 1]
    assert run("oracle", CASES["crossed_implied"])["has_infeasible"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_restatement(name):
    print(f"[singleton rows] {name}", flush=True)
    ref = run("oracle", CASES[name])
    got = run("gpu", CASES[name])
    for k in G.PLAN_FLAGS:
        assert got[k] == ref[k], k
    if ref["has_infeasible"]:
        return
    assert np.array_equal(got["keep_row_mask"], ref["keep_row_mask"])
    for k in ("new_l", "new_u"):
        assert np.array_equal(got[k].view(np.uint64), ref[k].view(np.uint64)), k
