"""HIP engine vs the CPU restatement on the same seeded inputs.

Integer / mask / map / tape results must be bit-identical.  Floating-point
results are compared bit-exactly as well (both sides use the same fixed
summation order and no FMA contraction); the objective constant, which
both sides sum in the same fixed tree, is also compared exactly.
"""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from backends import GpuBackend, OracleBackend

pytestmark = pytest.mark.gpu

RULES_READY = ["fixed_variables", "empty_columns", "empty_rows", "singleton_rows", "redundant_rows",
               "bound_strengthening", "singleton_columns", "dual_fix", "doubleton_equations", "parallel_rows",
               "parallel_columns"]


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    return a


def assert_same(got, exp, what):
    got, exp = np.asarray(got), np.asarray(exp)
    assert got.shape == exp.shape, f"{what}: shape {got.shape} vs {exp.shape}"
    if got.dtype == np.float64:
        # identical bits, treating +0 / -0 and NaN payloads as equal
        g, e = got.copy(), exp.copy()
        g[g == 0] = 0.0
        e[e == 0] = 0.0
        same = (bits(g) == bits(e)) | (np.isnan(g) & np.isnan(e))
    else:
        same = got == exp
    if not same.all():
        k = int(np.nonzero(~same)[0][0])
        raise AssertionError(f"{what}: {int((~same).sum())} mismatches, first at {k}: {got[k]!r} vs {exp[k]!r}")


def small_lp(seed):
    return G.generate_workload("small", seed=seed)


def test_transpose_matches_host():
    lp = G.generate_workload("small", seed=3)
    with G.DeviceLP(lp) as d:
        h = G.device_lp_to_host(d.lp)
    rp, ci, va = lp.transpose()
    assert_same(h.at[0], rp, "AT row_ptr")
    assert_same(h.at[1], ci, "AT col_idx")
    assert_same(h.at[2], va, "AT vals")


def test_transpose_zipf_medium():
    lp = G.generate_workload("medium", seed=0)
    print(f"[transpose] medium nnz={lp.nnz}", flush=True)
    with G.DeviceLP(lp) as d:
        h = G.device_lp_to_host(d.lp)
    rp, ci, va = lp.transpose()
    assert_same(h.at[0], rp, "AT row_ptr")
    assert_same(h.at[1], ci, "AT col_idx")
    assert_same(h.at[2], va, "AT vals")


def test_transpose_dense_and_long_columns():
    """Columns of every length class of the transpose's segment sorts: short
    (<= 32 entries), long (33..8192) and dense (> 8192, radix-sorted)."""
    rng = np.random.default_rng(5)
    m, n = 20000, 300
    rows, cols = [np.arange(m)], [np.zeros(m, dtype=np.int64)]          # dense column 0
    for j in range(1, 40):                                              # long columns
        r = rng.choice(m, size=int(rng.integers(33, 3000)), replace=False)
        rows.append(r)
        cols.append(np.full(r.size, j))
    for j in range(40, n):                                              # short columns
        r = rng.choice(m, size=int(rng.integers(0, 33)), replace=False)
        rows.append(r)
        cols.append(np.full(r.size, j))
    r, c = np.concatenate(rows), np.concatenate(cols)
    import scipy.sparse as sp
    A = sp.csr_matrix((rng.normal(size=r.size), (r, c)), shape=(m, n))
    A.sort_indices()
    lp = G.HostLP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy(), np.zeros(n),
                  np.zeros(n), np.ones(n), np.full(m, -np.inf), np.full(m, np.inf))
    with G.DeviceLP(lp) as d:
        h = G.device_lp_to_host(d.lp)
    rp, ci, va = lp.transpose()
    assert_same(h.at[0], rp, "AT row_ptr")
    assert_same(h.at[1], ci, "AT col_idx")
    assert_same(h.at[2], va, "AT vals")


def test_activity_dyadic_exact():
    """Mixed row lengths with dyadic data: sums are exact in any order
    (test_gpu_presolve_kernels.cpp:49-137)."""
    lengths = [0, 1, 8, 9, 31, 32, 33, 1024, 1025, 4097]
    rows, cols = 2100, 4097
    j = np.arange(cols)
    lower = np.where(j % 113 == 0, -np.inf, -(j % 7 + 1) * 0.5)
    upper = np.where(j % 127 == 0, np.inf, (j % 11 + 1) * 0.5)
    rp, ci, va = [0], [], []
    exp_lo, exp_hi = np.zeros(rows), np.zeros(rows)
    exp_ni, exp_pi = np.zeros(rows, np.int32), np.zeros(rows, np.int32)
    for i in range(rows):
        L = lengths[i % 10]
        jj = np.arange(L)
        a = np.where(jj % 71 == 0, 0.0, np.where(jj % 2 == 1, -0.5, 2.0))
        ci.extend(jj.tolist())
        va.extend(a.tolist())
        rp.append(len(ci))
        nz = a != 0
        lo = np.where(a >= 0, a * lower[jj], a * upper[jj])[nz]
        hi = np.where(a >= 0, a * upper[jj], a * lower[jj])[nz]
        exp_lo[i] = lo[np.isfinite(lo)].sum()
        exp_hi[i] = hi[np.isfinite(hi)].sum()
        exp_ni[i] = int((~np.isfinite(lo)).sum())
        exp_pi[i] = int((~np.isfinite(hi)).sum())
    lp = G.HostLP(rows, cols, np.asarray(rp, np.int32), np.asarray(ci, np.int32), np.asarray(va),
                  np.zeros(cols), lower, upper, np.zeros(rows), np.zeros(rows))
    import ctypes as C
    lib = G.load()
    with G.DeviceLP(lp) as d:
        outs = [G.DeviceArray(rows * 8), G.DeviceArray(rows * 8), G.DeviceArray(rows * 4), G.DeviceArray(rows * 4)]
        G.check(lib.gp_row_activity_summary(C.byref(d.lp.A), d.lp.l, d.lp.u, 1e-12, None, None,
                                            *[o.ptr for o in outs]))
        lo = G.to_host(outs[0].ptr, rows, np.float64)
        hi = G.to_host(outs[1].ptr, rows, np.float64)
        ni = G.to_host(outs[2].ptr, rows, np.int32)
        pi = G.to_host(outs[3].ptr, rows, np.int32)
    assert_same(lo, exp_lo, "min_fin")
    assert_same(hi, exp_hi, "max_fin")
    assert_same(ni, exp_ni, "min_ninf")
    assert_same(pi, exp_pi, "max_pinf")
    olo, ohi, oni, opi = O.activity_summary(lp, lower, upper, 1e-12)
    assert_same(olo, exp_lo, "oracle min_fin")
    assert_same(ohi, exp_hi, "oracle max_fin")


def random_state(lp, seed):
    rng = np.random.default_rng(seed)
    keep_row = (rng.random(lp.m) > 0.15).astype(np.uint8)
    keep_col = (rng.random(lp.n) > 0.15).astype(np.uint8)
    return keep_row, keep_col


@pytest.mark.parametrize("rule", RULES_READY)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rule_parity_random_state(rule, seed):
    print(f"[rule parity] {rule} seed={seed}", flush=True)
    lp = small_lp(seed)
    kr, kc = random_state(lp, 100 + seed)
    # batch-size gate of the doubleton rule off, so small LPs exercise it
    params = G.parity_params(record_postsolve_tape_cpu=1, doubleton_equations_min_selected_per_batch=1,
                             doubleton_equations_min_selected_ratio=0.0)
    ob, gb = OracleBackend(lp), GpuBackend(lp)
    try:
        for b in (ob, gb):
            b.set("keep_row_mask", kr)
            b.set("keep_col_mask", kc)
            b.refresh()
        for f in O.Plan._STATS:
            assert_same(gb.get(f), ob.get(f), f"stats {f}")
        for b in (ob, gb):
            b.apply(rule, params)
        for f in O.Plan._FIELDS:
            assert_same(gb.get(f), ob.get(f), f"{rule}: {f}")
        assert gb.flags() == ob.flags(), (gb.flags(), ob.flags())
        assert_same(np.array([gb.obj()]), np.array([ob.obj()]), "obj delta")
        tg, to = gb.tape(), ob.tape()
        for k in to:
            assert_same(tg[k], to[k], f"{rule}: tape {k}")
        ag, ao = gb.new_A(), ob.new_A()
        assert (ag is None) == (ao is None)
        if ao is not None:
            for x, y, k in zip(ag, ao, ("row_ptr", "col_idx", "vals")):
                assert_same(x, y, f"{rule}: new_A {k}")
    finally:
        gb.free()


def compare_pipeline(lp, params, label):
    print(f"[pipeline] {label} m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    ores = O.presolve(lp, params)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            rec = s.record()
            red = s.reduced()
            assert s.iterations == ores.iterations
            assert s.has_infeasible == ores.has_infeasible and s.has_unbounded == ores.has_unbounded
            assert s.has_change == ores.has_change
            assert (s.reduced_rows, s.reduced_cols, s.reduced_nnz) == (ores.m1, ores.n1, ores.reduced.nnz)
            assert_same(rec.row_red2org, ores.row_red2org, "row_red2org")
            assert_same(rec.col_red2org, ores.col_red2org, "col_red2org")
            assert_same(rec.fixed_idx, ores.fixed_idx, "fixed_idx")
            assert_same(rec.fixed_val, ores.fixed_val, "fixed_val")
            for k in ("types", "index_starts", "value_starts", "dual_modes", "indices", "vals"):
                assert_same(getattr(rec, k), getattr(ores, k), f"tape {k}")
            for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
                assert_same(getattr(red, k), getattr(ores.reduced, k), f"reduced {k}")
            for a, b, k in zip(red.at, ores.reduced.transpose(), ("AT row_ptr", "AT col_idx", "AT vals")):
                assert_same(a, b, k)
            assert_same(np.array([s.obj_constant_delta]), np.array([ores.obj_constant_delta]), "obj constant")
            return s, ores, rec
        except Exception:
            s.free()
            raise


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_pipeline_parity_small(seed):
    lp = small_lp(seed)
    s, ores, rec = compare_pipeline(lp, G.parity_params(), f"small seed={seed}")
    s.free()


def test_pipeline_parity_cpu_tape_records():
    lp = small_lp(7)
    s, ores, rec = compare_pipeline(lp, G.parity_params(record_postsolve_tape_cpu=1), "small cpu tape")
    s.free()


def test_pipeline_parity_medium():
    lp = G.generate_workload("medium", seed=0)
    s, ores, rec = compare_pipeline(lp, G.parity_params(), "medium seed=0")
    s.free()


def test_pipeline_parity_large10():
    """The bench recipe (integer data, Zipf columns) at 1/10 scale: 500k x 800k."""
    lp = G.generate_workload("large10", seed=0)
    s, ores, rec = compare_pipeline(lp, G.parity_params(), "large10 seed=0")
    s.free()


def test_postsolve_parity_small():
    lp = small_lp(5)
    params = G.parity_params(record_postsolve_tape_cpu=1)
    ores = O.presolve(lp, params)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(ores.n1)
    y = rng.standard_normal(ores.m1)
    z = rng.standard_normal(ores.n1)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            xg, yg, zg = G.postprocess(s, x, y, z, d)
        finally:
            s.free()
    xo, yo, zo = O.postsolve(ores, lp, x, y, z)
    assert_same(xg, xo, "postsolve x")
    assert_same(yg, yo, "postsolve y")
    assert_same(zg, zo, "postsolve z")


def test_public_pipeline_contract():
    """Identity run borrows the input, input unchanged, free twice is safe
    (test_public_pipeline.cpp:31-60)."""
    lp = small_lp(11)
    with G.DeviceLP(lp) as d:
        before = G.device_lp_to_host(d.lp)
        s = G.preprocess(d, G.parity_params(max_iters=0))
        assert s.owns_reduced_lp == 0 and s.has_change == 0
        assert s.reduced_lp.A.row_ptr == d.lp.A.row_ptr
        s.free()
        s.free()
        s = G.preprocess(d, G.parity_params())
        assert s.owns_reduced_lp == 1
        assert s.unsupported_stages == 0
        s.free()
        # reference defaults request folding (run; this LP has no symmetry to
        # fold) and the Startup rules (reported, not run)
        s = G.preprocess(d, G.default_params())
        assert s.unsupported_stages == 2 and s.folding_applied == 0 and s.folding.rounds >= 0
        s.free()
        s = G.preprocess(d, G.parity_params())
        assert s.preprocessing_seconds >= s.elapsed_seconds >= 0.0
        after = G.device_lp_to_host(d.lp)
        for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
            assert_same(getattr(after, k), getattr(before, k), f"input {k} unchanged")
        s.free()
        s.free()


def test_incremental_live_stats_match_full_refresh():
    """Live statistics updated from deleted lines (default) and re-swept in
    full give the same presolve, tape and reduced LP."""
    lp = G.generate_workload("medium", seed=1)
    with G.DeviceLP(lp) as d:
        out = []
        for inc in (0, 1):
            s = G.preprocess(d, G.parity_params(incremental_live_stats=inc, record_postsolve_tape_cpu=1))
            try:
                rec, red = s.record(), s.reduced()
                out.append((s.iterations, s.obj_constant_delta, rec, red))
            finally:
                s.free()
    (i0, o0, r0, d0), (i1, o1, r1, d1) = out
    assert i0 == i1
    assert_same(np.array([o1]), np.array([o0]), "obj constant")
    for k in ("types", "index_starts", "value_starts", "indices", "vals", "row_red2org", "col_red2org"):
        assert_same(getattr(r1, k), getattr(r0, k), f"record {k}")
    for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
        assert_same(getattr(d1, k), getattr(d0, k), f"reduced {k}")


@pytest.mark.parametrize("first,second,seed", [("bound_strengthening", "redundant_rows", 0),
                                               ("singleton_columns", "singleton_columns", 2),
                                               ("bound_strengthening", "bound_strengthening", 0)])
def test_rule_sequence_with_plan_rewrites(first, second, seed):
    """A caller of the public rule entry point rewrites the plan between two
    rule calls (bounds of a column the first rule left alone, a column turned
    back on).  The second call must see the rewritten plan: nothing cached by
    the first call may survive into it."""
    print(f"[rule sequence] {first} -> rewrite -> {second}", flush=True)
    lp = small_lp(4)
    params = G.parity_params(record_postsolve_tape_cpu=1)
    # 15% of the rows and 1% of the columns off; seeds where both calls reduce
    rng = np.random.default_rng(seed)
    kr = (rng.random(lp.m) > 0.15).astype(np.uint8)
    kc0 = (rng.random(lp.n) > 0.01).astype(np.uint8)
    ob, gb = OracleBackend(lp), GpuBackend(lp)
    try:
        for b in (ob, gb):
            b.set("keep_row_mask", kr)
            b.set("keep_col_mask", kc0)
            b.refresh()
            b.apply(first, params)
        for f in O.Plan._FIELDS:
            assert_same(gb.get(f), ob.get(f), f"{first}: {f}")
        l0, u0 = ob.get("new_l"), ob.get("new_u")
        kc = ob.get("keep_col_mask")
        same = np.nonzero((l0 == lp.l) & (u0 == lp.u) & (kc == 1))[0]
        off = np.nonzero(kc == 0)[0]
        rng = np.random.default_rng(5)
        pick = rng.choice(same, size=min(40, same.size), replace=False)
        l1 = l0.copy()
        # tighten finite lower bounds upward, leave infinite ones
        fin = np.isfinite(l1[pick]) & np.isfinite(u0[pick])
        l1[pick[fin]] = l1[pick[fin]] + 0.25 * (u0[pick[fin]] - l1[pick[fin]])
        kc1 = kc.copy()
        if off.size:
            kc1[off[: max(1, off.size // 3)]] = 1
        for b in (ob, gb):
            b.set("new_l", l1)
            b.set("keep_col_mask", kc1)
            b.refresh()
            b.apply(second, params)
        for f in O.Plan._FIELDS:
            assert_same(gb.get(f), ob.get(f), f"{second} after rewrite: {f}")
        assert gb.flags() == ob.flags(), (gb.flags(), ob.flags())
        tg, to = gb.tape(), ob.tape()
        for k in to:
            assert_same(tg[k], to[k], f"{second}: tape {k}")
    finally:
        gb.free()
