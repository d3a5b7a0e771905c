"""Structural Startup rules on the CPU restatement (oracle/oracle_startup.c):
the reference's known answers (cpp/tests/test_rule_column_domination.cpp,
test_rule_projected_auxiliaries.cpp) and what the mathematics fixes -- a
projected or dominated column never changes the optimum: the presolved LP
has the original's optimal objective (HiGHS), and the postsolved point is
feasible for the original LP and reaches that objective."""
import numpy as np
import pytest

import gpupre as G
import lp_check as H
import oracle as O
import startup_cases as S


def presolve(lp, params):
    r = O.presolve(lp, params)
    return r


def test_covering_reduction_and_primal_recovery():
    # test_exact_reduction_and_primal_recovery (test_rule_column_domination.cpp:135-174)
    lp = S.covering_lp()
    r = presolve(lp, S.covering_params())
    assert not r.has_infeasible and not r.has_unbounded
    assert (r.m1, r.n1, r.reduced.nnz) == (3, 4, 6)
    assert r.primal_only & G.PRIMAL_ONLY_COLUMN_DOMINATION
    assert sorted(r.fixed_idx.tolist()) == [3, 5] and list(r.fixed_val) == [0.0, 0.0]
    x, _, _ = O.postsolve(r, None, [1.0, 1.0, 1.0, 0.0], [0.0] * 3, [0.0] * 4)
    assert np.allclose(x, [1, 1, 1, 0, 0, 0], atol=1e-12)


def test_covering_rule_is_owned_by_startup():
    # test_rule_is_owned_by_startup (:176-188)
    r = presolve(S.covering_lp(), S.covering_params(enable_startup=0))
    assert (r.m1, r.n1, r.reduced.nnz) == (3, 6, 11)
    assert not r.primal_only


@pytest.mark.parametrize("variant", ["nonunit", "noncovering", "missing_witness", "nonbinary"])
def test_covering_fails_closed(variant):
    # assert_fail_closed (:190-200)
    ok, count = O.startup_analyze(S.covering_lp(variant), "column_domination", S.covering_params())
    assert not ok and count == 0


def test_covering_off_without_structure_rules():
    r = presolve(S.covering_lp(), S.covering_params(enable_structure_specific_rules=0))
    assert (r.m1, r.n1) == (3, 6)


def test_chained_projection_and_primal_recovery():
    # test_chained_projection_and_primal_recovery (test_rule_projected_auxiliaries.cpp:200-244)
    lp = S.pair_block_lp()
    r = presolve(lp, S.projected_params())
    assert not r.has_infeasible and not r.has_unbounded
    assert (r.m1, r.n1, r.reduced.nnz) == (0, 4, 0)
    assert r.primal_only & G.PRIMAL_ONLY_PROJECTED_AUXILIARY
    assert list(r.types) == [G.RED_FME_COL, G.RED_FME_COL]
    x, _, _ = O.postsolve(r, lp, [1.0, 1.0, 0.4, 0.2], np.zeros(0), np.zeros(4))
    assert np.allclose(x, [1.0, 1.0, 0.4, 0.2, 1.0, 0.1], atol=1e-12, rtol=0)


def test_projection_records():
    # The records of the pair block, written out: the two-row column 5
    # (rows 3 and 4, lower side from row 3: x5 >= -1 + x2/2 - x3/2 + x4,
    # floor 0) and then the product 4 (rows 0, 1, 2; x4 >= x0 + x1 - 1).
    r = presolve(S.pair_block_lp(), S.projected_params())
    idx = r.indices.tolist()
    assert idx[:r.index_starts[1]] == [2, 5, 3, 4, 3, 2, 1, 0, 1, 3, 2, 3, 4]
    assert idx[r.index_starts[1]:] == [3, 4, 0, 1, 2, 2, 2, 1, 0, 1, 2, 0, 1]
    assert r.vals.tolist() == [0.0, -1.0, 0.5, -0.5, 1.0, 0.0, -1.0, 1.0, 1.0]


def test_bounded_projection_fails_closed():
    # test_bounded_projection_fails_closed (:246-266)
    for lp in (S.pair_block_lp(3.0, 0.0), S.pair_block_lp(4.0, 0.25)):
        ok, count = O.startup_analyze(lp, "two_row_elimination", S.projected_params())
        assert not ok and count == 0


def test_mccormick_detector_is_exact_and_fail_closed():
    # test_mccormick_detector_is_exact_and_fail_closed (:268-292)
    p = S.projected_params()
    assert O.startup_analyze(S.mccormick_lp(), "zero_cost_products", p) == (True, 1)
    assert not O.startup_analyze(S.mccormick_lp(corrupt_upper_row=True), "zero_cost_products", p)[0]
    assert not O.startup_analyze(S.mccormick_lp(nonunit_endpoint=True), "zero_cost_products", p)[0]


def test_projection_rules_are_startup_only():
    # test_rules_are_startup_only (:294-307)
    r = presolve(S.pair_block_lp(), S.projected_params(enable_startup=0))
    assert (r.m1, r.n1, r.reduced.nnz) == (5, 6, 15) and len(r.types) == 0


def test_batch_matches_separate_applies():
    # test_batch_matches_separate_applies (:311-344): the shared plan equals
    # two materialized presolves, and so do the postsolved points
    lp = S.pair_block_lp()
    p = S.projected_params(zero_cost_product_variables_min_candidate_ratio=0.19)
    batch = presolve(lp, p)
    first = presolve(lp, S.projected_params(zero_cost_product_variables_min_candidate_ratio=0.19,
                                            enable_zero_cost_product_variables=0))
    second = presolve(first.reduced, S.projected_params(zero_cost_product_variables_min_candidate_ratio=0.19,
                                                        enable_two_row_variable_elimination=0))
    assert (first.m1, first.n1) == (3, 5)
    assert (batch.m1, batch.n1) == (second.m1, second.n1) == (0, 4)
    assert batch.reduced.nnz == second.reduced.nnz
    for f in ("c", "l", "u"):
        assert np.array_equal(getattr(batch.reduced, f), getattr(second.reduced, f))
    x = [1.0, 1.0, 0.4, 0.2]
    xb, _, _ = O.postsolve(batch, lp, x, np.zeros(0), np.zeros(4))
    x2, _, _ = O.postsolve(second, first.reduced, x, np.zeros(0), np.zeros(4))
    x1, _, _ = O.postsolve(first, lp, x2, np.zeros(first.m1), np.zeros(first.n1))
    assert np.array_equal(xb, x1)


def test_pending_masks_and_tape():
    # test_pending_masks_and_gpu_tape (:346-398): the rules against a plan that
    # already deletes row 0 and column 6 write the compact reference's records
    # with row ids shifted by the extra row, and delete every live line
    ref = presolve(S.pair_block_lp(), S.projected_params())
    lp = S.pending_lp()
    p = S.projected_params(zero_cost_product_variables_min_candidate_ratio=0.19)
    plan = O.Plan(lp)
    plan.set("keep_row_mask", [0, 1, 1, 1, 1, 1])
    plan.set("keep_col_mask", [1, 1, 1, 1, 1, 1, 0])
    plan.set("new_l", np.zeros(7))
    plan.set("new_u", [1, 1, 1, 1, 1, 1, 0])
    assert plan.startup("two_row_elimination", p) == (True, 1)
    assert plan.startup("zero_cost_products", p) == (True, 1)
    t = plan.tape()
    idx = t["indices"].copy()
    for k in range(len(t["types"])):
        b = t["index_starts"][k]
        idx[b + 2:b + 2 + idx[b]] -= 1
    assert idx.tolist() == ref.indices.tolist()
    assert t["vals"].tolist() == ref.vals.tolist()
    assert t["types"].tolist() == ref.types.tolist()
    assert plan.get("keep_row_mask").tolist() == [0] * 6
    assert plan.get("keep_col_mask").tolist() == [1, 1, 1, 1, 0, 0, 0]


# ------------------------------------------------------------ mathematical pins
def optimum_kept(lp, params):
    r = presolve(lp, params)
    assert not r.has_infeasible and not r.has_unbounded
    st, obj, *_ = H.solve(lp)
    if r.m1 == 0 and r.n1 == 0:
        x_red, obj_red = np.zeros(0), r.reduced.obj_constant
        st2 = 0
    else:
        st2, obj_red, x_red, *_ = H.solve(r.reduced)
    assert st == 0 and st2 == 0
    assert H.relative_gap(obj, obj_red) <= 1e-9
    x, _, _ = O.postsolve(r, lp, x_red, np.zeros(r.m1), np.zeros(r.n1))
    assert H.primal_residual(lp, x) <= 1e-7
    assert np.all(x >= lp.l - 1e-9) and np.all(x <= lp.u + 1e-9)
    assert H.relative_gap(float(lp.c @ x) + lp.obj_constant, obj) <= 1e-9
    return r


@pytest.mark.parametrize("seed", range(3))
def test_covering_keeps_the_optimum(seed):
    lp = S.planted_covering(seed, m=60, n=400, density=0.02)
    r = optimum_kept(lp, S.covering_params(max_iters=512, **{k: 1 for k in (
        "enable_infeasible_and_fixed_variables", "enable_empty_rows", "enable_empty_columns",
        "enable_singleton_rows", "enable_singleton_columns", "enable_dual_fix")}))
    assert r.primal_only & G.PRIMAL_ONLY_COLUMN_DOMINATION


@pytest.mark.parametrize("seed", range(4))
def test_projections_keep_the_optimum(seed):
    lp = S.planted_projections(seed, blocks=24)
    r = optimum_kept(lp, S.projected_params(max_iters=512, enable_bound_strengthening=1, enable_dual_fix=1,
                                            enable_singleton_rows=1, enable_empty_rows=1, enable_empty_columns=1))
    assert r.primal_only & G.PRIMAL_ONLY_PROJECTED_AUXILIARY
    assert (r.types == G.RED_FME_COL).sum() >= 10


def test_default_thresholds_hold_small_counts_back():
    # 24 blocks stay under the default 128-candidate floor
    lp = S.planted_projections(0, blocks=24)
    assert not O.startup_analyze(lp, "two_row_elimination", G.default_params())[0]
    assert not O.startup_analyze(lp, "zero_cost_products", G.default_params())[0]


@pytest.mark.parametrize("n", [6, 8])
def test_qap_reduction_matches_reference_formula(n):
    """The reference's QAP20 logs (results/raw/benchmarks/QAP20/03_GPU-Preprocessor/
    logs/*.log, e.g. tai60b: 26,355,720 x 12,963,600 -> 432,120 x 6,485,400 with
    26,359,200 nonzeros) follow 2n^3 + 2n rows, n^2 + (n^4 - n^2)/2 + n^2
    columns and n^2 (2n + 2) + 2n^4 nonzeros; the general rules alone give
    them on the generator's QAP LPs."""
    assert (2 * 60**3 + 2 * 60, 60**2 + (60**4 - 60**2) // 2 + 60**2,
            60**2 * (2 * 60 + 2) + 2 * 60**4) == (432120, 6485400, 26359200)
    lp = G.generate_qap(n, seed=1)
    assert (lp.m, lp.n) == (2 * n**4 + 2 * n**3 + n**2 + 2 * n, n**4 + n**2)
    r = O.presolve(lp, G.parity_params())
    assert (r.m1, r.n1, r.reduced.nnz) == (2 * n**3 + 2 * n, n * n + (n**4 - n * n) // 2 + n * n,
                                           n * n * (2 * n + 2) + 2 * n**4)
    on = O.presolve(lp, G.parity_params(enable_structure_specific_rules=1))
    assert (on.m1, on.n1, on.reduced.nnz) == (r.m1, r.n1, r.reduced.nnz) and on.primal_only == 0
