"""Known-answer programs: the reference's generated per-rule cases, built here
from the same recipes (cpp/tests/test_rule_*.cpp) and checked against the
expectations those tests state.  Each program takes a backend kind ("oracle"
or "gpu") so the CPU restatement and the HIP engine face identical cases."""
from __future__ import annotations

import numpy as np

import gpupre as G
from backends import make_backend

INF = np.inf
RED_SINGLETON_COLUMN_SUBSTITUTION = 2  # PostsolveReductionType::SingletonColumnSubstitution


def lp_from_columns(m, n, col_ptr, row_idx, vals, c=None):
    """HostLP whose A is the transpose of the given column lists (explicit
    zeros kept), rows in increasing column order."""
    col_ptr = np.asarray(col_ptr, np.int64)
    row_idx = np.asarray(row_idx, np.int32)
    vals = np.asarray(vals, np.float64)
    cols = np.repeat(np.arange(n, dtype=np.int32), np.diff(col_ptr))
    order = np.lexsort((cols, row_idx))
    rp = np.zeros(m + 1, np.int32)
    np.add.at(rp, row_idx + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    z = np.zeros
    return G.HostLP(m, n, rp, cols[order].astype(np.int32), vals[order], z(n) if c is None else np.asarray(c, float),
                    z(n), z(n), z(m), z(m))


def lp_from_rows(m, n, row_ptr, col_idx, vals):
    z = np.zeros
    return G.HostLP(m, n, np.asarray(row_ptr, np.int32), np.asarray(col_idx, np.int32),
                    np.asarray(vals, np.float64), z(n), z(n), z(n), z(m), z(m))


def run(kind, lp, plan, stats=None, params=None, calls=1):
    b = make_backend(kind, lp)
    for k, v in plan.items():
        b.set(k, v)
    for k, v in (stats or {}).items():
        b.set(k, v)
    b.set_live(0)
    return b


def finish(b):
    if hasattr(b, "free"):
        b.free()


def bits_equal(a, e):
    a = np.ascontiguousarray(np.asarray(a, np.float64)).view(np.uint64)
    e = np.ascontiguousarray(np.asarray(e, np.float64)).view(np.uint64)
    return a.shape == e.shape and bool((a == e).all())


# ---------------------------------------------------------------- singleton columns
def sc_eq_objective_bitwise(kind):
    """test_rule_singleton_columns.cpp:124-192"""
    lp = lp_from_rows(2, 3, [0, 2, 4], [0, 2, 1, 2], [3.0, 0.1, 7.0, 0.2])
    b = run(kind, lp, {"keep_row_mask": [1, 1], "keep_col_mask": [1, 1, 1],
                       "new_c": [100000000.1, -99999999.9, 0.3], "new_l": [-INF, -INF, 0.0],
                       "new_u": [INF, INF, 0.0], "new_AL": [5.0, 6.0], "new_AU": [5.0, 6.0]},
            {"singleton_column_mask": [1, 1, 0], "singleton_column_row": [0, 1, -1],
             "singleton_column_val": [3.0, 7.0, 0.0]})
    try:
        b.apply("singleton_columns", G.default_params())
        expected = 0.3
        expected += -(100000000.1 * 0.1 / 3.0)
        expected += -(-99999999.9 * 0.2 / 7.0)
        assert bits_equal(b.get("new_c")[2:], [expected])
    finally:
        finish(b)


def sc_long_column_objective_order(kind, m):
    """test_rule_singleton_columns.cpp:389-435"""
    n = m + 1
    ptr, idx, val = [0], [], []
    cost = np.zeros(n)
    cost[m] = 0.3
    expected = cost[m]
    support_row = np.full(n, -1, np.int32)
    support_val = np.zeros(n)
    single = np.ones(n, np.uint8)
    single[m] = 0
    for row in range(m):
        pivot, coefficient = (7.0, 0.2) if row % 2 else (3.0, 0.1)
        idx += [row, m]
        val += [pivot, coefficient]
        ptr.append(len(idx))
        cost[row] = -99999999.9 if row % 2 else 100000000.1
        expected += -(cost[row] * coefficient / pivot)
        support_row[row] = row
        support_val[row] = pivot
    lp = lp_from_rows(m, n, ptr, idx, val)
    lower, upper = np.full(n, -INF), np.full(n, INF)
    lower[m] = upper[m] = 0.0
    b = run(kind, lp, {"keep_row_mask": np.ones(m), "keep_col_mask": np.ones(n), "new_c": cost, "new_l": lower,
                       "new_u": upper, "new_AL": np.full(m, 5.0), "new_AU": np.full(m, 5.0)},
            {"singleton_column_mask": single, "singleton_column_row": support_row,
             "singleton_column_val": support_val})
    try:
        b.apply("singleton_columns", G.default_params())
        assert not b.flags()["has_infeasible"] and not b.flags()["has_unbounded"]
        assert bits_equal(b.get("new_c")[m:], [expected])
    finally:
        finish(b)


def sc_batched_tape(kind):
    """test_rule_singleton_columns.cpp:437-551 (two batches; the reference
    seeds the tape with one record first, here the tape starts empty)."""
    lengths = [1, 2, 1024, 1025, 2051, 3]
    m = len(lengths)
    ptr, idx, coeff, owner, rhs = [0], [], [], [], []
    cost, lower, upper, keep, single, single_row, single_val = [], [], [], [], [], [], []
    for row in range(m):
        first = len(idx)
        pivot_col = first + (lengths[row] - 1 if row % 2 else 0)
        pivot = -2.0 if row % 2 else 2.0
        owner.append(pivot_col)
        rhs.append(row + 1.0)
        for j in range(lengths[row]):
            col = first + j
            is_pivot = col == pivot_col
            a = pivot if is_pivot else (-0.0 if j % 5 == 0 else (j % 3 - 1) * 0.25)
            idx.append(col)
            coeff.append(a)
            cost.append(row + 0.5 if is_pivot else 0.125)
            lower.append(-INF if is_pivot else 0.0)
            upper.append(INF if is_pivot else 0.0)
            keep.append(1 if (is_pivot or j % 7 != 0) else 0)
            single.append(1 if is_pivot else 0)
            single_row.append(row if is_pivot else -1)
            single_val.append(pivot if is_pivot else 0.0)
        ptr.append(len(idx))
    n = len(idx)
    lp = lp_from_rows(m, n, ptr, idx, coeff)
    plan = {"keep_row_mask": np.ones(m), "keep_col_mask": keep, "new_c": cost, "new_l": lower, "new_u": upper,
            "new_AL": rhs, "new_AU": rhs}
    stats = {"singleton_column_mask": single, "singleton_column_row": single_row, "singleton_column_val": single_val}
    b = run(kind, lp, plan, stats)
    try:
        types, indices, vals, istarts, vstarts = [], [], [], [0], [0]
        for batch in range(2):
            for row in range(m):
                col = owner[row]
                support = [p for p in range(ptr[row], ptr[row + 1]) if idx[p] != col and keep[idx[p]]]
                if not support:
                    continue
                types.append(RED_SINGLETON_COLUMN_SUBSTITUTION)
                indices += [col, row, len(support)] + [idx[p] for p in support]
                vals += [single_val[col], rhs[row], lower[col], upper[col], cost[col], 1.0] + [coeff[p] for p in support]
                istarts.append(len(indices))
                vstarts.append(len(vals))
            b.apply("singleton_columns", G.default_params())
            fl = b.flags()
            assert not fl["has_infeasible"] and not fl["has_unbounded"]
            t = b.tape()
            assert list(t["types"]) == types
            assert list(t["index_starts"]) == istarts and list(t["value_starts"]) == vstarts
            assert list(t["indices"]) == indices
            assert bits_equal(t["vals"], vals)
            out_c, out_keep = b.get("new_c"), b.get("keep_col_mask")
            for row in range(m):
                support = [p for p in range(ptr[row], ptr[row + 1]) if idx[p] != owner[row] and keep[idx[p]]]
                assert out_keep[owner[row]] == (1 if not support else 0)
                for p in support:
                    col = idx[p]
                    exp = cost[col]
                    exp += -(cost[owner[row]] * coeff[p] / single_val[owner[row]])
                    assert bits_equal([out_c[col]], [exp])
            b.set("keep_row_mask", np.ones(m))
            b.set("keep_col_mask", keep)
            b.set("new_c", cost)
            b.set("new_AL", rhs)
            b.set("new_AU", rhs)
    finally:
        finish(b)


# ---------------------------------------------------------------- parallel columns
def pc_bounded_objective_prefilter(kind, ratio, gap_factor):
    """test_rule_parallel_columns.cpp:171-199"""
    m = 4097
    p = G.default_params()
    tol = max(p.zero_tol, p.bound_tol)
    gap = gap_factor * tol
    av = np.array([0.0 if i == 0 else 1.0 + (i % 3) for i in range(m)])
    lp = lp_from_columns(m, 2, [0, m, 2 * m], np.concatenate([np.arange(m), np.arange(m)]),
                         np.concatenate([av, ratio * av]))
    keep_rows = np.ones(m, np.uint8)
    keep_rows[0] = 0
    b = run(kind, lp, {"keep_row_mask": keep_rows, "keep_col_mask": [1, 1], "new_c": [2.0, ratio * 2.0 + gap],
                       "new_l": [-1.0, 2.0], "new_u": [3.0, 4.0], "new_AL": np.zeros(m), "new_AU": np.full(m, 10.0)})
    try:
        b.apply("parallel_columns", p)
        keep = b.get("keep_col_mask")
        t = b.tape()
        if gap > tol:
            assert list(keep) == [1, 1] and not b.flags()["has_change"] and len(t["types"]) == 0
        else:
            assert list(keep) == [1, 0] and b.flags()["has_change"] and len(t["types"]) == 1
            assert list(t["indices"]) == [1, 0] and t["vals"][0] == ratio
            lo, up = b.get("new_l"), b.get("new_u")
            assert lo[0] == -1.0 + ratio * (2.0 if ratio > 0 else 4.0)
            assert up[0] == 3.0 + ratio * (4.0 if ratio > 0 else 2.0)
    finally:
        finish(b)


def pc_large_group(kind, masked):
    """test_rule_parallel_columns.cpp:200-268"""
    m, n = 2048, 10
    ptr = np.arange(n + 1) * m
    rows = np.tile(np.arange(m), n)
    values = np.concatenate([(1.0 if j < 5 else 2.0) * (1.0 + (np.arange(m) % 4)) for j in range(n)])
    keep_rows, keep_cols = np.ones(m, np.uint8), np.ones(n, np.uint8)
    if masked:
        keep_rows[0] = 0
        keep_cols[n - 1] = 0
        for j in range(n):
            values[j * m] = 17.0 + j
    lp = lp_from_columns(m, n, ptr, rows, values)
    b = run(kind, lp, {"keep_row_mask": keep_rows, "keep_col_mask": keep_cols, "new_c": np.ones(n),
                       "new_l": np.zeros(n), "new_u": np.full(n, 10.0), "new_AL": np.zeros(m),
                       "new_AU": np.full(m, 10.0)})
    try:
        b.apply("parallel_columns", G.default_params())
        keep = b.get("keep_col_mask")
        assert keep[0] == 1 and keep[5] == 1
        assert all(keep[j] == 0 for j in range(1, n) if j != 5)
        up = b.get("new_u")
        assert up[0] == 50.0 and up[5] == (40.0 if masked else 50.0)
        assert len(b.tape()["types"]) == n - 2 - int(masked)
    finally:
        finish(b)


def pc_signature_case(kind, masked, scale_gate):
    """test_rule_parallel_columns.cpp:269-378 (plan/recovery expectations)"""
    m, n = 1300, 18
    ptr, rows, values = [0], [], []
    costs = np.ones(n)

    def append(length, start, scale, pattern):
        for r in range(length):
            rows.append(start + r)
            values.append(scale * ((1.0 + (r % 5)) if pattern == 0 else (1.0 + (r % 7) * 0.25)))
        ptr.append(len(rows))

    append(1100, 0, 1.0, 0); append(1100, 0, 4.0, 0); costs[1] = 4.0
    append(32, 0, 1.0, 1); append(32, 0, -4.0, 1); costs[3] = -4.0
    append(3, 20, 1.0, 0); append(3, 20, 2.0, 0); costs[5] = 2.0
    append(32, 0, 1.0, 0); append(1100, 0, 1.0, 0)
    append(1100, 0, 1.0, 1); append(0, 0, 1.0, 0)
    append(3, 0, 1e-12, 0)
    for _ in (11, 12):
        rows.extend([100, 101, 102])
        values.extend([1e-5, 1e300, 1e300])
        ptr.append(len(rows))
    for j in (13, 14, 15):
        rows.extend([200, 201, 202])
        values.extend([2.0 if j == 15 else 1.0, 1.0 if j == 13 else (-2.0 if j == 15 else -1.0),
                       1.0 if j == 13 else (-2.0 if j == 15 else -1.0)])
        costs[j] = 2.0 if j == 15 else 1.0
        ptr.append(len(rows))
    rows.extend([300, 301])
    values.extend([INF, 2.0])
    ptr.append(len(rows))
    append(64, 400, 1.0, 1)
    assert len(ptr) == n + 1
    values[ptr[2] + 2] = values[ptr[3] + 2] = 0.0
    keep_rows, keep_cols = np.ones(m, np.uint8), np.ones(n, np.uint8)
    keep_cols[7] = 0
    if masked:
        keep_rows[[0, 8, 1000]] = 0
        for j in (0, 1, 2, 3, 6, 7, 8):
            values[ptr[j]] = 19.0 + j
    lp = lp_from_columns(m, n, ptr, rows, values)
    kw = {}
    if scale_gate:
        kw = dict(parallel_columns_scale_guard_min_groups=2, parallel_columns_scale_guard_max_group_size=2)
    b = run(kind, lp, {"keep_row_mask": keep_rows, "keep_col_mask": keep_cols, "new_c": costs,
                       "new_l": np.full(n, -1.0), "new_u": np.full(n, 10.0), "new_AL": np.full(m, -2.0),
                       "new_AU": np.full(m, 20.0)})
    try:
        b.apply("parallel_columns", G.default_params(**kw))
        assert b.flags()["has_change"] and len(b.tape()["types"]) > 0
        keep = b.get("keep_col_mask")
        assert keep[7] == 0 and keep[9] == 1 and keep[10] == 1 and keep[12] == 0 and keep[15] == 0
        assert keep[1] == int(scale_gate) and keep[3] == int(scale_gate)
    finally:
        finish(b)


def pc_small_coefficients(kind, m, proportional):
    """test_rule_parallel_columns.cpp:506-532"""
    eps = 1.9877659e-8
    a0 = np.array([-1.0 if i == 0 else -eps for i in range(m)])
    a1 = np.array([1.0 if i == 0 else (eps if proportional else -eps) for i in range(m)])
    lp = lp_from_columns(m, 2, [0, m, 2 * m], np.concatenate([np.arange(m), np.arange(m)]), np.concatenate([a0, a1]))
    b = run(kind, lp, {"keep_row_mask": np.ones(m), "keep_col_mask": [1, 1], "new_c": [0.0, 0.0],
                       "new_l": [0.0, 0.0], "new_u": [INF, INF], "new_AL": np.zeros(m), "new_AU": np.full(m, INF)})
    try:
        b.apply("parallel_columns", G.default_params())
        keep = b.get("keep_col_mask")
        assert keep[0] == 1 and keep[1] == (0 if proportional else 1)
        assert bool(b.flags()["has_change"]) == proportional
        assert len(b.tape()["types"]) == (1 if proportional else 0)
        if proportional:
            assert b.get("new_l")[0] == -INF
    finally:
        finish(b)


PROGRAMS = [("sc_eq_objective_bitwise", sc_eq_objective_bitwise)]
PROGRAMS += [(f"sc_long_column_objective_order_m{m}", lambda k, m=m: sc_long_column_objective_order(k, m))
             for m in (9, 1025)]
PROGRAMS += [("sc_batched_tape", sc_batched_tape)]
PROGRAMS += [(f"pc_prefilter_r{r}_g{g}", lambda k, r=r, g=g: pc_bounded_objective_prefilter(k, r, g))
             for r in (2.0, -2.0) for g in (0.0, 0.25, 2.0)]
PROGRAMS += [(f"pc_large_group_masked{int(mk)}", lambda k, mk=mk: pc_large_group(k, mk)) for mk in (False, True)]
PROGRAMS += [(f"pc_signature_masked{int(mk)}_gate{int(g)}", lambda k, mk=mk, g=g: pc_signature_case(k, mk, g))
             for mk in (False, True) for g in (False, True)]
PROGRAMS += [(f"pc_small_coefficients_m{m}_p{int(p)}", lambda k, m=m, p=p: pc_small_coefficients(k, m, p))
             for m in (5, 2049) for p in (False, True)]
