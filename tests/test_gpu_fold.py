"""Folding on the GPU: gp_fold against the CPU restatement bit for bit
(classes, scales, folded LP) on the reference's folding test LPs and on LPs
made of identical copies; the folding stage of gp_preprocess (presolve on
the folded LP, postprocess through the unfold) against the restatement's
fold + presolve + postsolve + unfold; and the reference's pipeline tests
(cpp/tests/test_folding.cpp:150-316,339-387)."""
import numpy as np
import pytest

import fold_cases as F
import gpupre as G
import lp_check as H
import oracle as O

pytestmark = pytest.mark.gpu


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.dtype == np.float64:
        ok = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
    else:
        ok = a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} mismatches, first at {int(np.flatnonzero(~ok)[0])}"


def compare_fold(lp, tol=1e-8, label=""):
    ref = O.fold(lp, tol)
    with G.DeviceLP(lp) as d, G.Fold(d, tol) as f:
        r = f.r
        print(f"[fold] {label}: {lp.m} x {lp.n} -> applied={r.applied} {r.folded_rows} x {r.folded_cols} "
              f"rounds={r.rounds} ok={r.refinement_ok}", flush=True)
        assert bool(r.applied) == ref.applied, label
        assert (r.rounds, bool(r.refinement_ok)) == (ref.rounds, ref.refinement_ok), label
        assert (r.final_row_colors, r.final_col_colors) == (ref.mf, ref.nf), label
        if not ref.applied:
            assert (r.folded_rows, r.folded_cols) == (lp.m, lp.n)
            return ref, None
        rc, cc = f.classes()
        same(rc, ref.row_class, f"{label} row classes")
        same(cc, ref.col_class, f"{label} column classes")
        rs, cs = f.scales()
        same(rs, ref.row_scale, f"{label} row scales")
        same(cs, ref.col_scale, f"{label} column scales")
        got = f.folded()
        want = ref.folded
        for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
            same(getattr(got, k), getattr(want, k), f"{label} folded {k}")
        for a, b, k in zip(got.at, want.at, ("AT row_ptr", "AT col_idx", "AT vals")):
            same(a, b, f"{label} folded {k}")
        assert got.obj_constant == want.obj_constant
        return ref, got


def test_reference_folding_lps():
    compare_fold(F.symmetric_2x2(), label="symmetric 2x2")
    compare_fold(F.symmetric_2x2(vals=(1.0, 2.0, 2.0, 1.0), AL=(3.0, 3.0), AU=(3.0, 3.0), obj_constant=7.0),
                 label="equitable 2x2")
    compare_fold(F.unequal_sizes(), label="unequal sizes")
    compare_fold(F.symmetric_2x2(c=(2.0, 3.0), AL=(4.0, 5.0)), label="distinct metadata")
    compare_fold(F.diagonal_5x5(), label="rejected 5x5")
    for value in (0.0, -0.0, 1.25, 0.1, np.inf, -np.inf, 5e-324, 1.7976931348623157e308):
        for varied in (False, True):
            compare_fold(F.constant_columns(value, varied), label=f"constant {value} varied={varied}")


@pytest.mark.parametrize("seed", range(6))
def test_copies_fold_like_the_restatement(seed):
    lp, copies = F.symmetric_lp(seed, copies=3 + seed % 3, integer=seed % 2 == 1)
    ref, got = compare_fold(lp, label=f"copies seed {seed}")
    assert ref.applied


def test_large_symmetric_lp():
    lp, copies = F.symmetric_lp(42, copies=400, p=40, q=60, links=10)
    ref, got = compare_fold(lp, label="400 copies")
    assert ref.applied and got.n * copies == lp.n


def test_asymmetric_lp_is_passed_through():
    compare_fold(G.generate_workload("small", seed=0), label="small generator LP")


@pytest.mark.parametrize("seed", [0, 1])
def test_pipeline_matches_restatement(seed):
    lp, copies = F.symmetric_lp(10 + seed, copies=5)
    params = G.parity_params(enable_folding=1)
    ref = O.fold(lp, params.folding_tolerance)
    assert ref.applied
    ores = O.presolve(ref.folded, params)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            assert s.folding_applied and s.folding.applied
            assert (s.folding.folded_rows, s.folding.folded_cols) == (ref.mf, ref.nf)
            assert (s.original_rows, s.original_cols) == (lp.m, lp.n)
            assert (s.reduced_rows, s.reduced_cols, s.iterations) == (ores.m1, ores.n1, ores.iterations)
            red, rec = s.reduced(), s.record()
            for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
                same(getattr(red, k), getattr(ores.reduced, k), f"reduced {k}")
            for k in ("row_red2org", "col_red2org", "types", "indices", "vals"):
                same(getattr(rec, k), getattr(ores, k), f"record {k}")
            # a HiGHS solution of the reduced LP, postsolved and unfolded
            st, obj_red, xr, yr, zr = H.solve(ores.reduced)
            assert st == 0
            xo, yo, zo = G.postprocess(s, xr, yr, zr, d)
        finally:
            s.free()
    px, py, pz = O.postsolve(ores, ref.folded, xr, yr, zr)
    ux, uy, uz = ref.unfold(px, py, pz)
    same(xo, ux, "unfolded x")
    same(yo, uy, "unfolded y")
    same(zo, uz, "unfolded z")
    st, obj, *_ = H.solve(lp)
    assert st == 0
    assert H.relative_gap(float(lp.c @ xo) + lp.obj_constant, obj) <= 1e-6
    assert H.primal_residual(lp, xo) <= 1e-6


def test_reference_pipeline_equitable_case():
    # test_folding.cpp:192-262: folding before any presolve rule, then the
    # full presolve eliminating the folded LP
    lp = F.symmetric_2x2(vals=(1.0, 2.0, 2.0, 1.0), AL=(3.0, 3.0), AU=(3.0, 3.0), obj_constant=7.0)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, G.default_params(max_iters=0))
        try:
            assert s.folding_applied and (s.reduced_rows, s.reduced_cols) == (1, 1)
            red = s.reduced()
            assert red.obj_constant == 7.0 and abs(red.vals[0] - 3.0) < 1e-12 and abs(red.c[0] - 4.0) < 1e-12
            x, y, z = G.postprocess(s, [1.0], [4.0 / 3.0], [0.0], d)
        finally:
            s.free()
        vals = [1.0, 2.0, 2.0, 1.0]
        for j in range(2):
            assert abs(x[j] - 1.0) < 1e-12
            aty = vals[j] * y[0] + vals[2 + j] * y[1]
            assert abs(aty + z[j] - 2.0) < 1e-12 and abs(z[j]) < 1e-12
        assert abs(2 * (x[0] + x[1]) + 7 - 11) < 1e-12
        s = G.preprocess(d, G.default_params(max_iters=10))
        try:
            assert s.folding_applied and (s.reduced_rows, s.reduced_cols) == (0, 0)
            assert abs(s.reduced_lp.obj_constant - 11.0) < 1e-12
            x, y, z = G.postprocess(s, [], [], [], d)
        finally:
            s.free()
        assert all(abs(v - 1.0) < 1e-12 for v in x) and x[0] == x[1] and y[0] == y[1] and z[0] == z[1]


def test_presolve_switch_off_with_folding():
    # test_folding.cpp:339-387
    lp = F.symmetric_2x2()
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, G.default_params(enable_presolve=0, enable_folding=1))
        try:
            assert not s.presolve_enabled and s.folding_applied
            assert (s.reduced_rows, s.reduced_cols) == (1, 1) and s.iterations == 0
            x, y, z = G.postprocess(s, [2.0], [2.0], [0.0], d)
        finally:
            s.free()
        assert list(x) == [2.0, 2.0] and list(y) == [1.0, 1.0] and list(z) == [0.0, 0.0]
        s = G.preprocess(d, G.default_params(enable_presolve=0, enable_folding=0))
        try:
            assert not s.folding_applied and (s.reduced_rows, s.reduced_cols) == (2, 2)
            x, y, z = G.postprocess(s, [2.0, 2.0], [1.0, 1.0], [0.0, 0.0], d)
        finally:
            s.free()
        assert list(x) == [2.0, 2.0] and list(y) == [1.0, 1.0] and list(z) == [0.0, 0.0]
