"""Folding on the CPU: the restatement (oracle/oracle_fold.c) against the
reference's folding tests (cpp/tests/test_folding.cpp), and against what
the mathematics fixes: the classes of an LP made of identical copies merge
the copies (an equitable partition: every row of a class has the same sum
over every column class), and the folded LP has the original's optimal
objective (checked with HiGHS)."""
import numpy as np
import pytest

import fold_cases as F
import lp_check as H
import oracle as O

INF = np.inf


def test_identical_rows_and_columns_fold_to_one():
    # test_folding.cpp:98-122 and :389-405
    f = O.fold(F.symmetric_2x2(), 1e-8)
    assert f.applied and (f.mf, f.nf) == (1, 1)
    assert list(f.row_class) == [1, 1] and list(f.col_class) == [1, 1]
    assert list(f.row_scale) == [0.5, 0.5] and list(f.col_scale) == [0.5, 0.5]
    lp = f.folded
    assert (lp.m, lp.n, lp.nnz) == (1, 1, 1)
    assert lp.vals[0] == 2.0 and lp.c[0] == 4.0 and lp.AL[0] == 4.0 and lp.AU[0] == 4.0
    assert lp.l[0] == 0.0 and lp.u[0] == 10.0


def test_unfold_copies_x_and_scales_duals():
    # test_folding.cpp:124-148
    f = O.fold(F.symmetric_2x2(), 1e-8)
    x, y, z = f.unfold([4.0], [6.0], [8.0])
    assert list(x) == [4.0, 4.0] and list(y) == [3.0, 3.0] and list(z) == [4.0, 4.0]


def test_equitable_partition_values():
    # test_folding.cpp:194-220: A = [[1, 2], [2, 1]], sides 3 -> folded entry 3, cost 4, constant kept
    f = O.fold(F.symmetric_2x2(vals=(1.0, 2.0, 2.0, 1.0), AL=(3.0, 3.0), AU=(3.0, 3.0), obj_constant=7.0), 1e-8)
    assert f.applied and f.folded.vals[0] == 3.0 and f.folded.c[0] == 4.0 and f.folded.obj_constant == 7.0


def test_unequal_class_sizes():
    # test_folding.cpp:318-336
    f = O.fold(F.unequal_sizes(), 1e-8)
    assert f.applied and (f.mf, f.nf) == (1, 1)
    x, y, z = f.unfold([2.0], [2.0], [0.0])
    assert np.all(np.abs(x - 2) < 1e-12) and np.all(np.abs(y - 1) < 1e-12)
    assert np.all(np.abs(y[0] + y[1] + z - 2) < 1e-12)


def test_distinct_metadata_is_not_folded():
    # test_folding.cpp:291-303: distinct costs and sides make every class a singleton
    f = O.fold(F.symmetric_2x2(c=(2.0, 3.0), AL=(4.0, 5.0)), 1e-8)
    assert not f.applied and (f.mf, f.nf) == (2, 2)


def test_rejection_stops_after_one_round():
    # test_folding.cpp:411-437
    f = O.fold(F.diagonal_5x5(), 1e-8)
    assert not f.applied and not f.refinement_ok and f.rounds == 1 and (f.mf, f.nf) == (5, 5)


@pytest.mark.parametrize("value", [0.0, -0.0, 1.25, 0.1, INF, -INF, 5e-324, 1.7976931348623157e308])
@pytest.mark.parametrize("varied", [False, True])
def test_class_reductions_are_ordered_sums(value, varied):
    # test_folding.cpp:439-471: cost sums and bound / side averages equal the
    # left-to-right sums over the members, bit for bit
    lp = F.constant_columns(value, varied)
    f = O.fold(lp, 1e-8)
    assert f.applied and (f.mf, f.nf) == (1, 1)

    def ordered(v):
        s = 0.0
        for x in v:
            s += float(x)
        return s

    n = lp.n
    expect = [ordered(lp.c), ordered(lp.l) / n, ordered(lp.u) / n, ordered(lp.AL) / 3, INF]
    got = [f.folded.c[0], f.folded.l[0], f.folded.u[0], f.folded.AL[0], f.folded.AU[0]]
    assert np.array_equal(np.array(got).view(np.uint64), np.array(expect).view(np.uint64)) or \
        all((np.isnan(a) and np.isnan(b)) for a, b in zip(got, expect))


def equitable(lp, rc, cc):
    """every row of a row class has the same sum over every column class"""
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    sums = {}
    for i, j, a in zip(rows, lp.col_idx, lp.vals):
        sums[(i, cc[j])] = sums.get((i, cc[j]), 0.0) + a
    per_class = {}
    for (i, k), s in sums.items():
        per_class.setdefault((rc[i], k), {})[i] = s
    for (r, k), d in per_class.items():
        members = np.flatnonzero(rc == r)
        vals = [d.get(i, 0.0) for i in members]
        if max(vals) - min(vals) > 1e-9:
            return False
    return True


@pytest.mark.parametrize("seed", range(4))
def test_copies_merge_and_the_optimum_is_kept(seed):
    lp, copies = F.symmetric_lp(seed, copies=4)
    f = O.fold(lp, 1e-8)
    assert f.applied
    # each class holds one line of every copy (4 lines), the link rows aside
    assert f.nf * copies == lp.n
    assert equitable(lp, f.row_class, f.col_class)
    st, obj, *_ = H.solve(lp)
    st2, obj2, x2, *_ = H.solve(f.folded)
    assert st == 0 and st2 == 0
    assert H.relative_gap(obj, obj2) <= 1e-9
    # the unfolded optimum of the folded LP is feasible and optimal for the original
    x, _, _ = f.unfold(x2, np.zeros(f.mf), np.zeros(f.nf))
    assert H.primal_residual(lp, x) <= 1e-7
    assert H.relative_gap(float(lp.c @ x) + lp.obj_constant, obj) <= 1e-9


def test_asymmetric_lp_is_not_folded():
    import gpupre as G

    lp = G.generate_workload("small", seed=1)
    f = O.fold(lp, 1e-8)
    assert not f.applied
