"""The single-row doubleton path (doubleton_equations_scan off,
rule_doubleton_equations.cu:1948-2128) on the CPU restatement.

Pinned three ways:
* the reference's own test (test_rule_doubleton_equations.cpp:43-87, in
  tests/golden/reference_kats.json as "doubleton_eliminates_row");
* hand-worked cases whose expected values are computed here by hand (row
  choice, one elimination per call, pivot and fill-in rejection, mapped
  bounds, the rewritten A with an inserted and a cancelled keep entry, the
  tape record, the infeasible case);
* the mathematics: whole presolves with the single-row path keep the optimum
  (HiGHS on the reduced LP, postsolved, against HiGHS on the original).
"""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from backends import OracleBackend
from lp_check import primal_residual, relative_gap, solve

INF = np.inf


def single_row_params(**kw):
    return G.parity_params(doubleton_equations_scan=0, **kw)


def backend(A, c, l, u, AL, AU):
    lp = G.HostLP.from_dense(np.asarray(A, float), np.asarray(c, float), np.asarray(l, float),
                             np.asarray(u, float), np.asarray(AL, float), np.asarray(AU, float))
    b = OracleBackend(lp)
    b.refresh()
    return b


def test_lowest_row_eliminated_one_per_call():
    # r0: x0 + 2 x1 = 4 ; r1: x2 + x3 = 3 ; r2: x0 + x2 + x3 <= 10
    A = [[1, 2, 0, 0], [0, 0, 1, 1], [1, 0, 1, 1]]
    b = backend(A, c=[1, 3, 0, 0], l=[0, 0, 0, 0], u=[5, 10, 9, 9], AL=[4, 3, -INF], AU=[4, 3, 10])
    b.apply("doubleton_equations", single_row_params(doubleton_equations_max_fill_in_proxy=10))
    # x1 is the only column with one live entry: it is eliminated, x0 kept;
    # x1 = 2 - x0 / 2 in [0, 10]  ->  x0 in [-16, 4], so x0 in [0, 4]
    assert list(b.get("keep_row_mask")) == [0, 1, 1]
    assert list(b.get("keep_col_mask")) == [1, 0, 1, 1]
    assert list(b.get("new_l")) == [0.0, 0.0, 0.0, 0.0]
    assert list(b.get("new_u")) == [4.0, 10.0, 9.0, 9.0]
    # c0 = 1 + 3 * (-1/2); objective constant 3 * 2
    assert list(b.get("new_c")) == [-0.5, 3.0, 0.0, 0.0]
    assert b.obj() == 6.0
    # r0's sides lose 2 * 2 (the row is deleted); the others are unchanged
    assert list(b.get("new_AL")) == [0.0, 3.0, -INF] and list(b.get("new_AU")) == [0.0, 3.0, 10.0]
    t = b.tape()
    assert list(t["types"]) == [G.RED_DOUBLETON_EQUATION] and list(t["indices"]) == [1, 0, 0, 0]
    assert list(t["vals"]) == [2.0, 1.0, 4.0, 0.0, 10.0, 0.0, 5.0, 0.0, 4.0, 3.0, 0.0, 0.0]
    # the rewritten A keeps every entry (no row holds x1 besides r0)
    rp, ci, va = b.new_A()
    assert list(rp) == [0, 2, 4, 7] and list(ci) == [0, 1, 2, 3, 0, 2, 3]
    # the second call takes r1
    b.set_flag("has_change", 0)
    b.apply("doubleton_equations", single_row_params(doubleton_equations_max_fill_in_proxy=10))
    assert list(b.get("keep_row_mask")) == [0, 0, 1]


def test_rewritten_A_inserts_and_cancels_the_kept_entry():
    # r0: x0 - x1 = 1 (x1 eliminated: x1 = x0 - 1); r1 holds x1 but not x0,
    # r2 holds both with a_x0 = -a_x1, so x0's entry cancels
    A = [[1, -1, 0], [0, 2, 1], [3, -3, 1]]
    b = backend(A, c=[0, 0, 1], l=[0, -INF, 0], u=[10, INF, 10], AL=[1, -INF, -INF], AU=[1, 8, 9])
    b.apply("doubleton_equations", single_row_params())
    # col_nnz: x0 2, x1 3 -> neither is 1; |1/-1| integral both ways -> the
    # column with fewer live entries (x0) is eliminated: x0 = 1 + x1
    assert list(b.get("keep_col_mask")) == [0, 1, 1]
    rp, ci, va = b.new_A()
    # r0 as it was; r1 unchanged (no x0); r2: x1 entry -3 + 1 * 3 = 0 dropped
    assert list(rp) == [0, 2, 4, 6]
    assert list(ci) == [0, 1, 1, 2, 0, 2] and list(va) == [1.0, -1.0, 2.0, 1.0, 3.0, 1.0]
    # r2's sides lose 3 * 1
    assert b.get("new_AU")[2] == 6.0 and b.get("new_AL")[2] == -INF
    # x1 = x0 - 1 with x0 in [0, 10]: x1 in [-1, 9]
    assert list(b.get("new_l"))[1] == -1.0 and list(b.get("new_u"))[1] == 9.0


def test_rewritten_A_inserts_the_kept_entry_in_column_order():
    # r0: x0 + x1 = 2, x0 in two rows and x1 in three: the ratio is integral
    # both ways, so the column with fewer live entries (x0) is eliminated:
    # x0 = 2 - x1.  r1 holds x0 but not x1: x1 enters r1 with -1 * 1 between
    # x0 and x2; x0's own entry stays (its column is dead).
    A = [[1, 1, 0], [1, 0, 1], [0, 1, 1], [0, 1, 0]]
    b = backend(A, c=[1, 0, 0], l=[0, 0, 0], u=[5, 5, 5], AL=[2, -INF, -INF, -INF], AU=[2, 5, 5, 5])
    b.apply("doubleton_equations", single_row_params())
    assert list(b.get("keep_col_mask")) == [0, 1, 1]
    rp, ci, va = b.new_A()
    assert list(rp) == [0, 2, 5, 7, 8]
    assert list(ci) == [0, 1, 0, 1, 2, 1, 2, 1]
    assert list(va) == [1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0]
    # r1's sides lose 1 * 2; x1 = 2 - x0 with x0 in [0, 5]: x1 in [-3, 2] -> [0, 2]
    assert b.get("new_AU")[1] == 3.0
    assert b.get("new_l")[1] == 0.0 and b.get("new_u")[1] == 2.0
    # c1 = 0 + 1 * (-1); objective constant 1 * 2
    assert b.get("new_c")[1] == -1.0 and b.obj() == 2.0


def test_pivot_ratio_and_fill_in_reject_rows():
    # r0: x0 + 2000 x1 = 1 (ratio 2000 > 1e3: rejected); r1: x2 + x3 = 1
    A = [[1, 2000, 0, 0], [0, 0, 1, 1]]
    b = backend(A, c=[0] * 4, l=[0] * 4, u=[1] * 4, AL=[1, 1], AU=[1, 1])
    b.apply("doubleton_equations", single_row_params())
    assert list(b.get("keep_row_mask")) == [1, 0]
    # fill-in: eliminating x0 would add x1 to 3 rows; with a proxy limit of 1
    # only the row whose elimination adds nothing passes
    A = [[1, 1, 0, 0], [1, 0, 1, 0], [1, 0, 0, 1], [1, 0, 0, 0], [0, 0, 1, 1]]
    dense = np.asarray(A, float)
    b = backend(dense, c=[0] * 4, l=[0] * 4, u=[5] * 4, AL=[1, -INF, -INF, -INF, 2], AU=[1, 9, 9, 9, 2])
    b.apply("doubleton_equations", single_row_params(doubleton_equations_max_fill_in_proxy=1))
    # r0: x1 (one entry) eliminated, kept x0: no fill -> r0 is taken
    assert list(b.get("keep_row_mask")) == [0, 1, 1, 1, 1]


def test_infeasible_keep_bounds():
    # x0 + x1 = 10 with x0, x1 in [0, 1]: x1 = 10 - x0 needs x0 >= 9
    A = [[1, 1], [1, 0]]
    b = backend(A, c=[0, 0], l=[0, 0], u=[1, 1], AL=[10, -INF], AU=[10, 5])
    b.apply("doubleton_equations", single_row_params())
    assert b.flags()["has_infeasible"] == 1
    assert list(b.get("keep_row_mask")) == [1, 1] and list(b.get("keep_col_mask")) == [1, 1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_single_row_presolve_keeps_the_optimum(seed):
    lp = G.generate_workload("small", seed=seed)
    st0, obj0, *_ = solve(lp)
    assert st0 == 0
    res = O.presolve(lp, single_row_params())
    assert not res.has_infeasible
    assert any(t == G.RED_DOUBLETON_EQUATION for t in res.types), "no single-row elimination ran"
    st, obj_r, xr, yr, zr = solve(res.reduced)
    assert st == 0
    x, y, z = O.postsolve(res, lp, xr, yr, zr)
    obj = float(np.dot(lp.c, x)) + lp.obj_constant
    assert relative_gap(obj, obj0) <= 1e-6, (obj, obj0)
    assert primal_residual(lp, x) <= 1e-6
