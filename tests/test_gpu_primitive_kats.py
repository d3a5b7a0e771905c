"""The reference's primitive known answers on the HIP engine through the C
ABI: FixedVariable tape with the long-double objective
(cpp/tests/test_fixed_variable_tape.cpp), the parallel-row hash closed form
(cpp/tests/test_gpu_presolve_kernels.cpp:139-179), incremental live
statistics against full sweeps (:181-310) and the stable-filter compaction of
a 32768 x 9000 LP with long rows (cpp/tests/test_gpu_direct_port.cpp:71-147)."""
import ctypes as C

import numpy as np
import pytest

import gpupre as G
import oracle as O
from test_gpu_parity import assert_same
from test_primitive_kats import (FT_AT, FT_COSTS, FT_EXCLUDED, FT_FIXED, FT_ROWS, FT_VALUES, compaction_fixture,
                                 hash_expected, hash_fixture)

pytestmark = pytest.mark.gpu


def dev(a, dt):
    return G.DeviceArray.from_host(np.asarray(a, dt))


def device_csr(rows, cols, rp, ci, va):
    d = [dev(rp, G.NP_OFFSET), dev(ci, np.int32), dev(va, np.float64)]
    return G.GpCsr(rows, cols, len(ci), d[0].ptr, d[1].ptr, d[2].ptr), d


def plan_tape(plan):
    lib = G.load()
    r, i, v = C.c_int64(), C.c_int64(), C.c_int64()
    G.check(lib.gp_plan_tape_sizes(C.byref(plan), C.byref(r), C.byref(i), C.byref(v)))
    t = dict(types=np.empty(r.value, np.int32), index_starts=np.empty(r.value + 1, G.NP_OFFSET),
             value_starts=np.empty(r.value + 1, G.NP_OFFSET), dual_modes=np.empty(r.value, np.uint8),
             indices=np.empty(i.value, np.int32), vals=np.empty(v.value, np.float64))
    if r.value:
        G.check(lib.gp_plan_tape_copy(C.byref(plan), *[t[k].ctypes.data_as(C.c_void_p) for k in
                                                      ("types", "index_starts", "value_starts", "dual_modes",
                                                       "indices", "vals")]))
    return t


def test_fixed_variable_tape_known_answer():
    print("[primitive kat] fixed-variable tape", flush=True)
    lib = G.load()
    rows, cols, rp, ci, va = FT_AT
    at, keep_at = device_csr(rows, cols, rp, ci, va)
    # the plan's LP: A = transpose of the fixture's AT (3 rows, 5 columns)
    dense = np.zeros((cols, rows))
    for j in range(rows):
        for p in range(rp[j], rp[j + 1]):
            dense[ci[p], j] = va[p]
    host = G.HostLP.from_dense(dense, np.zeros(rows), np.zeros(rows), np.ones(rows), np.zeros(cols), np.ones(cols))
    arrays = [dev(FT_FIXED, np.uint8), dev(FT_EXCLUDED, np.uint8), dev(FT_VALUES, np.float64),
              dev(FT_ROWS, np.uint8), dev(FT_COSTS, np.float64)]
    ptrs = [a.ptr for a in arrays]
    ref_obj, ref_tape = O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS)
    with G.DeviceLP(host) as d:
        for compute in (1, 0):
            plan = G.GpPlan()
            G.check(lib.gp_plan_create(C.byref(d.lp), C.byref(plan)))
            try:
                G.check(lib.gp_plan_enable_tape(C.byref(plan)))
                obj = C.c_double(-1.0)
                G.check(lib.gp_fixed_variable_tape(C.byref(plan), *ptrs, C.byref(at), rows, compute, C.byref(obj)))
                assert obj.value == (1.0 if compute else 0.0)
                t = plan_tape(plan)
                assert list(t["indices"]) == [0, 0, 1, 2, 2, 0, 2]
                assert list(t["index_starts"]) == [0, 2, 4, 7]
                for k in ref_tape:
                    assert_same(t[k], ref_tape[k], f"tape {k}")
            finally:
                lib.gp_plan_destroy(C.byref(plan))
    for compute, want in ((1, 1.0), (0, 0.0)):
        obj = C.c_double(-1.0)
        G.check(lib.gp_fixed_variable_tape(None, *ptrs, C.byref(at), rows, compute, C.byref(obj)))
        assert obj.value == want
    assert ref_obj == 1.0
    none = dev([0] * 5, np.uint8)
    obj = C.c_double(-1.0)
    G.check(lib.gp_fixed_variable_tape(None, none.ptr, *ptrs[1:], C.byref(at), rows, 1, C.byref(obj)))
    assert obj.value == 0.0
    G.check(lib.gp_fixed_variable_tape(None, None, None, None, None, None, C.byref(at), 0, 1, C.byref(obj)))
    assert obj.value == 0.0


def test_extended_objective_random_sums():
    """The x87 emulation against the CPU's long double on random products:
    magnitudes over 40 binades, cancellations, signed zeros and subnormals."""
    lib = G.load()
    rng = np.random.default_rng(3)
    n = 4000
    c = rng.standard_normal(n) * 10.0 ** rng.integers(-20, 20, n)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-20, 20, n)
    c[5::50] = 1e16
    c[6::50] = -1e16
    v[5::50] = v[6::50] = 1.0
    c[7::61] = -0.0
    v[8::67] = 5e-324
    c[8::67] = 0.5
    fixed = (rng.random(n) < 0.7).astype(np.uint8)
    at_rp = np.zeros(n + 1, G.NP_OFFSET)
    at, keep_at = device_csr(n, 1, at_rp, [], [])
    keep_row = dev([1], np.uint8)
    arrays = [dev(fixed, np.uint8), dev(v, np.float64), dev(c, np.float64)]
    obj = C.c_double()
    G.check(lib.gp_fixed_variable_tape(None, arrays[0].ptr, None, arrays[1].ptr, keep_row.ptr, arrays[2].ptr,
                                       C.byref(at), n, 1, C.byref(obj)))
    ref, _ = O.fixed_variable_tape((n, 1, at_rp, [], []), fixed, None, v, [1], c, with_tape=False)
    assert np.float64(obj.value).view(np.uint64) == np.float64(ref).view(np.uint64), (obj.value, ref)


@pytest.mark.parametrize("rows", [10, 2100])
def test_parallel_row_hash_closed_form(rows):
    print(f"[primitive kat] row hash rows={rows}", flush=True)
    ptr, idx, vals, keep = hash_fixture()
    exp = hash_expected(ptr, idx, vals, keep, rows)
    A, keep_a = device_csr(rows, 4097, ptr[:rows + 1], idx[:ptr[rows]], vals[:ptr[rows]])
    kr = dev(keep[:rows], np.uint8)
    out = G.DeviceArray(8 * rows)
    G.check(G.load().gp_parallel_row_hashes(C.byref(A), kr.ptr, None, 1e-12, out.ptr))
    got = G.to_host(out.ptr, rows, np.uint64)
    assert np.array_equal(got, exp)


def test_incremental_live_statistics_match_full_sweeps():
    """test_gpu_presolve_kernels.cpp:181-310: a 5 x 5 matrix with an explicit
    zero, masks shrinking through four states; the incremental update equals
    a full refresh each time, and a row coming back to life is rejected
    without touching the live total."""
    print("[primitive kat] incremental live stats", flush=True)
    lib = G.load()
    dense_rp = [0, 3, 5, 6, 8, 8]
    dense_ci = [0, 1, 3, 1, 2, 2, 3, 4]
    dense_va = [2, 0, 4, 5, 6, 7, 8, 9]
    host = G.HostLP(5, 5, np.asarray(dense_rp, G.NP_OFFSET), np.asarray(dense_ci, np.int32),
                    np.asarray(dense_va, np.float64), np.zeros(5), np.zeros(5), np.ones(5), np.zeros(5), np.ones(5))
    from backends import GpuBackend
    full, delta = GpuBackend(host, tape=False), GpuBackend(host, tape=False)
    try:
        for b in (full, delta):
            b.refresh()
        old_rows, old_cols = dev([1] * 5, np.uint8), dev([1] * 5, np.uint8)

        def compare():
            assert full.stats.live_nnz == delta.stats.live_nnz
            for f in full._STATS:
                assert_same(delta.get(f), full.get(f), f)

        compare()
        row_states = [[1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 0, 0, 0, 1], [1, 0, 0, 0, 0]]
        col_states = [[1, 1, 1, 1, 1], [1, 1, 1, 0, 1], [1, 1, 0, 0, 1], [1, 0, 0, 0, 1]]
        for rs, cs in zip(row_states, col_states):
            for b in (full, delta):
                b.set("keep_row_mask", rs)
                b.set("keep_col_mask", cs)
            ok = C.c_int32(-1)
            G.check(lib.gp_update_live_stats(C.byref(delta.stats), C.byref(delta.dlp.lp), C.byref(delta.plan),
                                             old_rows.ptr, old_cols.ptr, C.byref(ok)))
            assert ok.value == 1, "incremental update rejected a deletion"
            full.refresh()
            compare()
            G.to_device(old_rows.ptr, np.asarray(rs, np.uint8))
            G.to_device(old_cols.ptr, np.asarray(cs, np.uint8))
        before = delta.stats.live_nnz
        reactivated = list(row_states[-1])
        reactivated[1] = 1
        delta.set("keep_row_mask", reactivated)
        ok = C.c_int32(-1)
        G.check(lib.gp_update_live_stats(C.byref(delta.stats), C.byref(delta.dlp.lp), C.byref(delta.plan),
                                         old_rows.ptr, old_cols.ptr, C.byref(ok)))
        assert ok.value == 0, "reactivation was accepted"
        assert delta.stats.live_nnz == before
    finally:
        full.free()
        delta.free()


def test_compaction_stable_filter_long_rows():
    print("[primitive kat] compaction 32768 x 9000", flush=True)
    lp, params, e = compaction_fixture()
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            assert not s.has_infeasible and not s.has_unbounded
            rec, red = s.record(), s.reduced()
        finally:
            s.free()
    assert_same(rec.row_red2org, e["row_map"], "row map")
    assert_same(rec.col_red2org, e["col_map"], "col map")
    assert_same(red.row_ptr, e["row_ptr"], "row_ptr")
    assert_same(red.col_idx, e["col_idx"], "col_idx")
    assert np.array_equal(red.vals.view(np.uint64), e["vals"].view(np.uint64))


@pytest.mark.parametrize("n,end_bit,distinct", [(1, 64, 0), (2, 64, 0), (4095, 64, 0), (4096, 23, 0),
                                                (4097, 64, 3), (100_003, 64, 0), (100_003, 64, 17),
                                                (1_000_000, 40, 1000), (3_000_001, 24, 0)])
def test_radix_sort_pairs_is_a_stable_sort(n, end_bit, distinct):
    """Keys on bits [0, end_bit) -- uniform, or drawn from `distinct` values
    to make long runs of equal keys -- with values recording the input
    order: the device sort equals numpy's stable argsort, bits above end_bit
    ignored."""
    print(f"[radix sort] n={n} end_bit={end_bit} distinct={distinct}", flush=True)
    rng = np.random.default_rng(n + end_bit)
    if distinct:
        pool = rng.integers(0, 1 << 63, size=distinct, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        keys = pool[rng.integers(0, distinct, size=n)]
    else:
        keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n,
                                                                                             dtype=np.uint64)
    vals = np.arange(n, dtype=np.int32)
    mask = np.uint64((1 << end_bit) - 1) if end_bit < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    dk, dv = G.DeviceArray.from_host(keys), G.DeviceArray.from_host(vals)
    G.check(G.load().gp_radix_sort_pairs(dk.ptr, dv.ptr, n, end_bit))
    gk, gv = G.to_host(dk.ptr, n, np.uint64), G.to_host(dv.ptr, n, np.int32)
    assert np.array_equal(gv, order.astype(np.int32))
    assert np.array_equal(gk, keys[order])
