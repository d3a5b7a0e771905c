"""Scheduler known-answer programs: the cases of the reference's deferred
schedule test (cpp/tests/test_deferred_schedule.cpp), restated on the round
trace this engine records (phase, live nnz before / after, next phase),
the summary counters, the reduced LP and postsolve.  Each program runs on the
CPU restatement ("oracle") or the HIP engine ("gpu").  The reference's rule-level
trace lines (rule invocation counts, apply boundaries) have no counterpart in
this trace and are not checked."""
from __future__ import annotations

import numpy as np

import gpupre as G
import oracle as O

PHASE = {0: "cleanup", 1: "propagation", 2: "aggregation", 3: "parallel", 4: "end"}


def model(n, rows, al, au, lower=0.0, upper=10.0, c=None):
    """Model(n, rows, al, au, lower, upper) of test_deferred_schedule.cpp:25-45."""
    m = len(rows)
    A = np.asarray(rows, np.float64).reshape(m, n)
    cc = np.zeros(n) if c is None else np.asarray(c, np.float64)
    return G.HostLP.from_dense(A, cc, np.full(n, float(lower)), np.full(n, float(upper)),
                               np.asarray(al, np.float64), np.asarray(au, np.float64))


def parameters(**kw):
    """parameters() of test_deferred_schedule.cpp:46-53."""
    base = dict(enable_startup=0, bound_strengthening_max_bound_only_rounds=1, max_iters=12, collect_trace=1)
    base.update(kw)
    return G.parity_params(**base)


class Run:
    """One presolve on a backend: counters, trace, reduced LP, postsolve."""

    def __init__(self, kind, lp, params):
        self.kind, self.lp = kind, lp
        if kind == "oracle":
            r = O.presolve(lp, params)
            self.res = r
            self.iterations, self.stopped = r.iterations, r.stopped_at_iteration_limit
            self.rows, self.cols, self.nnz = r.m1, r.n1, r.reduced.nnz
            self.reduced = r.reduced
            self.trace = [(PHASE[t.phase], t.nnz_before, t.nnz_after, PHASE[t.next_phase]) for t in r.trace]
            self.types = r.types
            self.indices, self.vals = r.indices, r.vals
            self.obj_delta = r.obj_constant_delta
            self.infeasible, self.unbounded = r.has_infeasible, r.has_unbounded
        else:
            self.d = G.DeviceLP(lp)
            s = G.preprocess(self.d, params)
            self.s = s
            self.iterations, self.stopped = s.iterations, s.stopped_at_iteration_limit
            self.rows, self.cols, self.nnz = s.reduced_rows, s.reduced_cols, s.reduced_nnz
            self.reduced = s.reduced()
            self.trace = [(PHASE[t.phase], t.nnz_before, t.nnz_after, PHASE[t.next_phase]) for t in s.trace()]
            rec = s.record()
            self.types, self.indices, self.vals = rec.types, rec.indices, rec.vals
            self.obj_delta = s.obj_constant_delta
            self.infeasible, self.unbounded = s.has_infeasible, s.has_unbounded
        assert not self.infeasible and not self.unbounded

    def phases(self):
        return [t[0] for t in self.trace]

    def count(self, phase):
        return sum(1 for t in self.trace if t[0] == phase)

    def has(self, phase, before, after, nxt):
        return (phase, before, after, nxt) in self.trace

    def postsolve(self, x, y=None, z=None):
        x = np.asarray(x, np.float64)
        y = np.zeros(self.rows) if y is None else np.asarray(y, np.float64)
        z = np.zeros(self.cols) if z is None else np.asarray(z, np.float64)
        if self.kind == "oracle":
            return O.postsolve(self.res, self.lp, x, y, z)[0]
        return G.postprocess(self.s, x, y, z, self.d)[0]

    def close(self):
        if self.kind == "gpu":
            self.s.free()
            self.d.free()


def cleanup_follows_startup_at_iteration_limit(kind):
    """test_deferred_schedule.cpp:90-106"""
    lp = model(3, [[1, -1, -1]], [0], [0], 0, np.inf, c=[1, 0, 0])
    r = Run(kind, lp, parameters(enable_startup=1, max_iters=1))
    try:
        assert r.iterations == 1
        assert (r.rows, r.cols, r.nnz) == (0, 0, 0)
        assert r.phases() == ["cleanup", "cleanup", "cleanup"]
        assert np.all(r.postsolve(np.zeros(0)) == 0)
    finally:
        r.close()


def order(kind):
    """test_deferred_schedule.cpp:107-122"""
    lp = model(3, [[1, 2, 3], [2, -1, 1], [1, 1, -2]], [0, 0, 0], [0, 0, 0], -np.inf, np.inf)
    r = Run(kind, lp, parameters(max_iters=3))
    try:
        assert r.phases() == ["cleanup", "propagation", "aggregation", "parallel"]
        assert r.iterations == 3
        assert not r.stopped
    finally:
        r.close()


def cleanup_dependency_repeats_without_budget(kind, budget):
    """test_deferred_schedule.cpp:123-145"""
    lp = model(3, [[1, 0, 0], [1, 1, 0], [0, 1, 1]], [1, 2, 2], [1, 2, 2], 0, 5)
    p = parameters(enable_startup=1, max_iters=budget, enable_bound_strengthening=0,
                   enable_infeasible_and_redundant_rows=0, enable_dual_fix=0, enable_doubleton_equations=0,
                   enable_parallel_rows=0, enable_parallel_columns=0, enable_singleton_columns=0)
    r = Run(kind, lp, p)
    try:
        assert (r.cols, r.rows) == (0, 0)
        x = r.postsolve(np.ones(0))
        assert np.all(np.abs(x - 1) < 1e-12)
    finally:
        r.close()


def interleaved_progress_flags(kind):
    """test_deferred_schedule.cpp:146-166"""
    lp = model(2, [[1, 1], [1, -1]], [-np.inf, 0], [100, 0])
    p = parameters(max_iters=1, enable_infeasible_and_fixed_variables=0, enable_empty_columns=0,
                   enable_empty_rows=0, enable_singleton_rows=0, enable_singleton_columns=0,
                   bound_strengthening_bound_only_nnz_round_budget=1)
    r = Run(kind, lp, p)
    try:
        assert (r.rows, r.cols, r.nnz) == (1, 2, 2)
        assert list(r.reduced.u) == [10, 10]
        assert list(r.reduced.l) == [0, 0]
    finally:
        r.close()


def stage_progress_advances_after_no_change(kind, removed):
    """test_deferred_schedule.cpp:167-196"""
    pairs = (removed - 3) // 2 if removed % 2 else removed // 2
    rows = [[1, 1, 0] for _ in range(pairs)]
    if removed % 2:
        rows.append([1, 1, 1])
    al, au = [-np.inf] * len(rows), [100.0] * len(rows)
    rows.append([1, -1, 0])
    al.append(0)
    au.append(0)
    lp = model(3, rows, al, au)
    p = parameters(max_iters=4, enable_infeasible_and_fixed_variables=0, enable_empty_columns=0,
                   enable_empty_rows=0, enable_singleton_rows=0, enable_singleton_columns=0,
                   enable_doubleton_equations=0, enable_parallel_rows=0, enable_parallel_columns=0,
                   bound_strengthening_bound_only_nnz_round_budget=1)
    r = Run(kind, lp, p)
    try:
        repeats = removed >= 100
        assert r.iterations == (4 if repeats else 3) and r.nnz == 2
        assert r.count("propagation") == (2 if repeats else 1)
        assert r.has("propagation", removed + 2, 2, "cleanup" if repeats else "aggregation")
        assert r.has("propagation", 2, 2, "aggregation") == repeats
        assert r.count("aggregation") == 1
        assert r.count("parallel") == 1
    finally:
        r.close()


def parallel_progress_does_not_credit_earlier_phases(kind):
    """test_deferred_schedule.cpp:197-210"""
    rows = [[0.0] * 74 for _ in range(9)]
    for i in range(5):
        for j in range(20):
            rows[i][j] = 1 + (i * 7 + j * 3 + j * i) % 19
    for j in range(20, 70):
        rows[5][j] = rows[7][j] = rows[8][j] = 1
    rows[6][70] = rows[6][71] = 1
    al, au = [-np.inf] * 9, [5.0] * 9
    au[5] = 1000
    al[6] = au[6] = 1
    au[7] = au[8] = 15
    lp = model(74, rows, al, au)
    p = parameters(enable_infeasible_and_fixed_variables=0, enable_empty_columns=0, enable_empty_rows=0,
                   enable_singleton_rows=0, enable_singleton_columns=0, enable_bound_strengthening=0,
                   enable_dual_fix=0, enable_parallel_columns=0, doubleton_equations_min_selected_per_batch=1)
    r = Run(kind, lp, p)
    try:
        assert lp.nnz == 252 and r.nnz == 150
        assert r.iterations == 3 and not r.stopped
        assert r.count("propagation") == 1 and r.count("aggregation") == 1 and r.count("parallel") == 1
        assert r.has("propagation", 252, 202, "aggregation")
        assert r.has("aggregation", 202, 200, "parallel")
        assert r.has("parallel", 200, 150, "end")
    finally:
        r.close()


def single_pass_row_side_update(kind):
    """test_deferred_schedule.cpp:211-227"""
    lp = model(2, [[1, 1], [1, -1]], [0, 0], [15, 0])
    p = parameters(max_iters=1, enable_infeasible_and_fixed_variables=0, enable_empty_columns=0,
                   enable_empty_rows=0, enable_singleton_rows=0, enable_singleton_columns=0,
                   bound_strengthening_bound_only_nnz_round_budget=1)
    r = Run(kind, lp, p)
    try:
        assert (r.rows, r.cols, r.nnz) == (2, 2, 4)
        assert r.reduced.AL[0] == -np.inf and r.reduced.AL[1] == 0
    finally:
        r.close()


def carry_and_doubleton_equations(kind, max_iters):
    """test_deferred_schedule.cpp:228-257"""
    lp = model(3, [[1, 1, 0], [0, 1, 1], [1, 0, 1]], [2, 2, 2], [2, 2, 2])
    r = Run(kind, lp, parameters(max_iters=max_iters, doubleton_equations_min_selected_per_batch=1))
    try:
        assert r.stopped == (max_iters == 2)
        assert (r.nnz, r.rows, r.cols) == (4, 2, 2)
        i = r.phases().index("aggregation")
        assert r.trace[i] == ("aggregation", 6, 4, "parallel")
        assert "cleanup" not in r.phases()[i + 1:]
        x = r.postsolve(np.ones(r.cols))
        assert np.all(np.abs(x - 1) < 1e-9)
    finally:
        r.close()


def aggregation_iteration_limit(kind):
    """test_deferred_schedule.cpp:258-270"""
    lp = model(3, [[1, 1, 0], [0, 1, 1], [1, 0, 1]], [2, 2, 2], [2, 2, 2])
    r = Run(kind, lp, parameters(max_iters=1, doubleton_equations_min_selected_per_batch=1))
    try:
        assert r.iterations == 1 and r.stopped
        assert (r.rows, r.cols, r.nnz) == (3, 3, 6)
        assert "aggregation" not in r.phases()
    finally:
        r.close()


def parallel(kind):
    """test_deferred_schedule.cpp:271-284"""
    lp = model(3, [[1, 1, 2], [2, 2, 4]], [2, 4], [8, 16])
    p = parameters(enable_dual_fix=0, enable_bound_strengthening=0, enable_infeasible_and_redundant_rows=0,
                   enable_doubleton_equations=0)
    r = Run(kind, lp, p)
    try:
        i = r.phases().index("parallel")
        assert "cleanup" not in r.phases()[i + 1:]
        assert (r.rows, r.cols, r.nnz) == (1, 1, 1)
    finally:
        r.close()


def progress(kind, residual_nnz):
    """test_deferred_schedule.cpp:285-309"""
    rows = [[0.0] * 6 for _ in range(6)]
    for i in range(5):
        for j in range(5):
            if i != j:
                rows[i][j] = 1 + i * 5 + j
    rows[0][1] = 0
    if residual_nnz == 18:
        rows[0][2] = 0
    rows[5][5] = 1
    al, au = [0.0] * 6, [100.0] * 6
    al[5], au[5] = 1, 2
    lp = model(6, rows, al, au)
    p = parameters(enable_singleton_columns=0, enable_doubleton_equations=0, enable_dual_fix=0,
                   enable_bound_strengthening=0, enable_infeasible_and_redundant_rows=0, enable_parallel_rows=0,
                   enable_parallel_columns=0)
    r = Run(kind, lp, p)
    try:
        assert r.nnz == residual_nnz
        assert r.count("propagation") == 1
        assert r.count("cleanup") == 3
        assert r.has("cleanup", residual_nnz, residual_nnz, "cleanup")
        assert r.has("cleanup", residual_nnz, residual_nnz, "propagation")
        assert r.has("propagation", residual_nnz, residual_nnz, "aggregation")
        assert r.count("aggregation") == 1
        assert r.iterations == 3
    finally:
        r.close()


def limits_and_empty(kind):
    """test_deferred_schedule.cpp:310-315"""
    lp = model(1, [[1]], [2], [2])
    r = Run(kind, lp, parameters(max_iters=0))
    try:
        assert r.nnz == 1 and r.iterations == 0
    finally:
        r.close()
    r = Run(kind, lp, parameters(max_iters=10, max_time=1e-12))
    try:
        assert r.iterations == 0
    finally:
        r.close()


def small_cleanup_cascade_after_startup_budget(kind):
    """test_deferred_schedule.cpp:316-351"""
    background, n = 10, 13
    rows = [[0.0] * n for _ in range(background)]
    for i in range(background):
        for j in range(background):
            rows[i][j] = 1 + (i * 7 + j * 3 + i * j) % 19
    al, au = [-np.inf] * background, [np.inf] * background
    for k in range(3):
        row = [0.0] * n
        row[background + k] = 1
        if k:
            row[background + k - 1] = 1
        rows.append(row)
        al.append(2 if k else 1)
        au.append(2 if k else 1)
    lp = model(n, rows, al, au)
    p = parameters(enable_startup=1, max_iters=1, enable_singleton_columns=0,
                   bound_strengthening_max_bound_only_rounds=64,
                   bound_strengthening_bound_only_nnz_round_budget=104 * 3)
    r = Run(kind, lp, p)
    try:
        assert lp.nnz == 105 and r.iterations == 1
        assert (r.rows, r.cols, r.nnz) == (background, background, 100)
        assert r.count("cleanup") == 5
        assert r.has("cleanup", 105, 104, "cleanup")
        assert r.has("cleanup", 100, 100, "cleanup")
        assert r.has("cleanup", 100, 100, "propagation")
        assert r.count("propagation") == 0
        x = r.postsolve(np.zeros(background), np.zeros(background), np.zeros(background))
        assert np.all(x[:background] == 0)
        assert np.all(np.abs(x[background:] - 1) < 1e-12)
    finally:
        r.close()


def wide_singleton_recovery(kind, startup, coefficient_sign, variable_sign):
    """test_deferred_schedule.cpp:352-377"""
    n = 4096
    row = [float(-coefficient_sign)] * n
    row[-1] = float(coefficient_sign)
    cost = [0.0] * n
    cost[-1] = float(variable_sign)
    lp = model(n, [row], [0], [0], 0 if variable_sign > 0 else -np.inf, np.inf if variable_sign > 0 else 0,
               c=cost)
    r = Run(kind, lp, parameters(enable_startup=int(startup)))
    try:
        assert (r.rows, r.cols) == (0, 0)
        assert int(np.sum(r.types == G.RED_SINGLETON_COLUMN_SUBSTITUTION)) == 1
        assert r.indices.size < 4 * n and r.vals.size < 8 * n
        assert np.all(r.postsolve(np.zeros(0)) == 0)
        assert r.obj_delta == 0
    finally:
        r.close()


def fixed_bound_budget(kind, limit, redundant_rows, row_side_updates, initial_singleton):
    """test_deferred_schedule.cpp:378-440"""
    n = 8
    rows, al, au = [], [], []
    for j in range(1, n):
        row = [0.0] * n
        row[j - 1], row[j] = -1, 1
        rows.append(row)
        al.append(-100 if row_side_updates else -np.inf)
        au.append(0)
    if redundant_rows:
        for j in range(1, n):
            row = [0.0] * n
            row[0] = row[j] = 1
            rows.append(row)
            al.append(-np.inf)
            au.append(100)
    if initial_singleton:
        row = [0.0] * n
        row[0] = 1
        rows.append(row)
        al.append(-np.inf)
        au.append(1)
    lp = model(n, rows, al, au)
    if not initial_singleton:
        lp.u[0] = 1.0
    chain_nnz = 14 * (2 if redundant_rows else 1)
    p = parameters(enable_infeasible_and_fixed_variables=0, enable_empty_columns=0, enable_empty_rows=0,
                   enable_singleton_rows=int(initial_singleton), enable_singleton_columns=0, enable_dual_fix=0,
                   enable_doubleton_equations=0, enable_parallel_rows=0, enable_parallel_columns=0,
                   enable_infeasible_and_redundant_rows=int(redundant_rows or row_side_updates),
                   bound_strengthening_max_bound_only_rounds=64,
                   bound_strengthening_bound_only_nnz_round_budget=chain_nnz * limit)
    initial_nnz = chain_nnz + (1 if initial_singleton else 0)
    actual_limit = max(1, chain_nnz * limit // initial_nnz)
    r = Run(kind, lp, p)
    try:
        expected_calls = actual_limit + (1 if redundant_rows else 0)
        assert (r.rows, r.cols, r.nnz) == (7, n, 14)
        assert list(r.reduced.u) == [1.0 if j <= expected_calls else 10.0 for j in range(n)]
        assert r.count("propagation") == expected_calls
        if row_side_updates:
            assert np.all(np.asarray(r.reduced.AL) == -np.inf)
        if initial_singleton:
            assert r.has("cleanup", 15, 14, "cleanup")
    finally:
        r.close()


PROGRAMS = [("cleanup_follows_startup_at_iteration_limit", cleanup_follows_startup_at_iteration_limit),
            ("order", order)]
PROGRAMS += [(f"fixed_bound_budget_{a}_{int(b)}{int(c)}{int(d)}", lambda k, a=a, b=b, c=c, d=d:
              fixed_bound_budget(k, a, b, c, d))
             for a, b, c, d in [(1, False, False, False), (2, False, False, False), (4, False, False, False),
                                (2, True, False, False), (2, False, True, False), (2, False, False, True)]]
PROGRAMS += [(f"cleanup_dependency_budget{b}", lambda k, b=b: cleanup_dependency_repeats_without_budget(k, b))
             for b in (1, 2, 3, 4)]
PROGRAMS += [("interleaved_progress_flags", interleaved_progress_flags)]
PROGRAMS += [(f"stage_progress_removed{r}", lambda k, r=r: stage_progress_advances_after_no_change(k, r))
             for r in (99, 100, 102)]
PROGRAMS += [("parallel_progress_no_credit", parallel_progress_does_not_credit_earlier_phases),
             ("single_pass_row_side_update", single_pass_row_side_update)]
PROGRAMS += [(f"carry_doubleton_iters{i}", lambda k, i=i: carry_and_doubleton_equations(k, i)) for i in (12, 2)]
PROGRAMS += [("aggregation_iteration_limit", aggregation_iteration_limit), ("parallel", parallel)]
PROGRAMS += [(f"progress_{r}", lambda k, r=r: progress(k, r)) for r in (19, 18)]
PROGRAMS += [("limits_and_empty", limits_and_empty),
             ("small_cleanup_cascade_after_startup_budget", small_cleanup_cascade_after_startup_budget)]
PROGRAMS += [(f"wide_singleton_s{int(s)}_c{c}_v{v}", lambda k, s=s, c=c, v=v: wide_singleton_recovery(k, s, c, v))
             for s in (False, True) for c in (-1, 1) for v in (-1, 1)]
