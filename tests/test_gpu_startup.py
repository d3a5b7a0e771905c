"""Structural Startup rules on the GPU (gp_startup_rule and the Startup stage
of gp_preprocess) against the CPU restatement bit for bit: the reference's
known answers (cpp/tests/test_rule_column_domination.cpp,
test_rule_projected_auxiliaries.cpp), the fail-closed variants, analyses
against a pending plan, and seeded LPs with planted covering, two-row and
product structure (with decoys) through the whole pipeline and postsolve."""
import numpy as np
import pytest

import gpupre as G
import oracle as O
import startup_cases as S
from backends import GpuBackend, OracleBackend, gpu_startup_analyze
from test_gpu_parity import assert_same, compare_pipeline

pytestmark = pytest.mark.gpu


def pipeline(lp, params, label, x=None):
    s, ores, rec = compare_pipeline(lp, params, label)
    try:
        assert s.primal_only_reductions == ores.primal_only, label
        if x is None:
            rng = np.random.default_rng(3)
            x = rng.uniform(0.0, 1.0, ores.n1)
        with G.DeviceLP(lp) as d:
            xg, yg, zg = G.postprocess(s, x, np.zeros(ores.m1), np.zeros(ores.n1), d)
        xo, yo, zo = O.postsolve(ores, lp, x, np.zeros(ores.m1), np.zeros(ores.n1))
        assert_same(xg, xo, f"{label} postsolve x")
        assert_same(yg, yo, f"{label} postsolve y")
        assert_same(zg, zo, f"{label} postsolve z")
        return ores, xg
    finally:
        s.free()


def test_covering_known_answer():
    ores, x = pipeline(S.covering_lp(), S.covering_params(), "covering", x=np.array([1.0, 1.0, 1.0, 0.0]))
    assert (ores.m1, ores.n1, ores.reduced.nnz) == (3, 4, 6)
    assert sorted(ores.fixed_idx.tolist()) == [3, 5]
    assert ores.primal_only & G.PRIMAL_ONLY_COLUMN_DOMINATION
    assert np.allclose(x, [1, 1, 1, 0, 0, 0], atol=1e-12)
    ores, _ = pipeline(S.covering_lp(), S.covering_params(enable_startup=0), "covering, no Startup")
    assert (ores.m1, ores.n1, ores.reduced.nnz) == (3, 6, 11)


@pytest.mark.parametrize("variant", ["nonunit", "noncovering", "missing_witness", "nonbinary"])
def test_covering_fails_closed(variant):
    got = gpu_startup_analyze(S.covering_lp(variant), "column_domination", S.covering_params())
    assert got == (False, 0)


def test_projection_known_answers():
    lp = S.pair_block_lp()
    ores, x = pipeline(lp, S.projected_params(), "pair block", x=np.array([1.0, 1.0, 0.4, 0.2]))
    assert (ores.m1, ores.n1, ores.reduced.nnz) == (0, 4, 0)
    assert ores.types.tolist() == [G.RED_FME_COL, G.RED_FME_COL]
    assert np.allclose(x, [1.0, 1.0, 0.4, 0.2, 1.0, 0.1], atol=1e-12, rtol=0)
    ores, _ = pipeline(lp, S.projected_params(enable_startup=0), "pair block, no Startup")
    assert (ores.m1, ores.n1, ores.reduced.nnz) == (5, 6, 15)
    p = S.projected_params()
    for bad in (S.pair_block_lp(3.0, 0.0), S.pair_block_lp(4.0, 0.25)):
        assert gpu_startup_analyze(bad, "two_row_elimination", p) == (False, 0)
    assert gpu_startup_analyze(S.mccormick_lp(), "zero_cost_products", p) == (True, 1)
    assert not gpu_startup_analyze(S.mccormick_lp(corrupt_upper_row=True), "zero_cost_products", p)[0]
    assert not gpu_startup_analyze(S.mccormick_lp(nonunit_endpoint=True), "zero_cost_products", p)[0]


def test_pending_plan_matches_restatement():
    lp = S.pending_lp()
    p = S.projected_params(zero_cost_product_variables_min_candidate_ratio=0.19)
    states = []
    for B in (OracleBackend, GpuBackend):
        b = B(lp)
        b.set("keep_row_mask", [0, 1, 1, 1, 1, 1])
        b.set("keep_col_mask", [1, 1, 1, 1, 1, 1, 0])
        b.set("new_l", np.zeros(7))
        b.set("new_u", [1, 1, 1, 1, 1, 1, 0])
        got = (b.startup("two_row_elimination", p), b.startup("zero_cost_products", p))
        states.append((got, b.tape(), b.get("keep_row_mask"), b.get("keep_col_mask"), b.flags()))
        if B is GpuBackend:
            b.free()
    (og, ot, orow, ocol, of), (gg, gt, grow, gcol, gf) = states
    assert og == gg == ((True, 1), (True, 1))
    for k in ot:
        assert_same(gt[k], ot[k], f"pending tape {k}")
    assert_same(grow, orow, "keep_row")
    assert_same(gcol, ocol, "keep_col")
    assert gf == of


@pytest.mark.parametrize("seed", range(4))
def test_planted_projections_pipeline(seed):
    lp = S.planted_projections(seed, blocks=40)
    params = S.projected_params(max_iters=512, enable_bound_strengthening=1, enable_dual_fix=1,
                                enable_singleton_rows=1, enable_empty_rows=1, enable_empty_columns=1)
    ores, _ = pipeline(lp, params, f"planted projections seed {seed}")
    assert ores.primal_only & G.PRIMAL_ONLY_PROJECTED_AUXILIARY


@pytest.mark.parametrize("seed", range(3))
def test_planted_covering_pipeline(seed):
    lp = S.planted_covering(seed, m=200, n=3000, density=0.003)
    ores, _ = pipeline(lp, S.covering_params(max_iters=512, enable_empty_columns=1, enable_singleton_rows=1),
                       f"planted covering seed {seed}")
    assert ores.primal_only & G.PRIMAL_ONLY_COLUMN_DOMINATION


def test_default_parameters_at_scale():
    """Reference defaults (all general rules, default thresholds, the
    structural rules on) on LPs large enough to pass the thresholds."""
    params = G.parity_params(enable_structure_specific_rules=1)
    lp = S.planted_projections(11, blocks=1500, extra_rows=800)
    ores, _ = pipeline(lp, params, "planted projections, 1500 blocks")
    assert ores.primal_only & G.PRIMAL_ONLY_PROJECTED_AUXILIARY
    lp = S.planted_covering(12, m=600, n=9000, density=0.002)
    ores, _ = pipeline(lp, params, "planted covering, 9000 columns")
    assert ores.primal_only & G.PRIMAL_ONLY_COLUMN_DOMINATION


def test_structural_rules_leave_generator_lps_alone():
    """The generator LP has no such structure: with the structural rules on,
    the reduction equals the one with them off, on both sides."""
    lp = G.generate_workload("small", seed=2)
    on, _ = pipeline(lp, G.parity_params(enable_structure_specific_rules=1), "small, structural rules on")
    off = O.presolve(lp, G.parity_params())
    assert (on.m1, on.n1, on.reduced.nnz) == (off.m1, off.n1, off.reduced.nnz)
    assert on.primal_only == 0
