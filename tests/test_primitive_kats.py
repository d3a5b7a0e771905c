"""Known answers of the reference's primitive tests on the CPU restatement:
the FixedVariable tape and its long-double objective
(cpp/tests/test_fixed_variable_tape.cpp), the parallel-row hash closed form
(cpp/tests/test_gpu_presolve_kernels.cpp:139-179) and the stable-filter
compaction oracle (cpp/tests/test_gpu_direct_port.cpp:71-147)."""
import numpy as np
import pytest

import oracle as O

# cpp/tests/test_fixed_variable_tape.cpp:35-44
FT_AT = (5, 3, [0, 2, 3, 5, 7, 8], [0, 1, 2, 0, 2, 1, 2, 0], [-0.0, 8, 2, 3, 4, 5, 6, 7])
FT_FIXED = [2, 1, 1, 1, 0]
FT_EXCLUDED = [0, 0, 0, 1, 0]
FT_ROWS = [2, 0, 1]
FT_VALUES = [1, 1, 1, 999, -0.0]
FT_COSTS = [1e16, 1, -1e16, 99, 7]


def test_fixed_variable_tape_known_answer():
    obj, tape = O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS)
    # the stable cancellation 1e16 + 1 - 1e16 = 1 of the long-double sum
    assert obj == 1.0
    assert list(tape["indices"]) == [0, 0, 1, 2, 2, 0, 2]
    assert list(tape["index_starts"]) == [0, 2, 4, 7]
    assert list(tape["types"]) == [0, 0, 0] and list(tape["dual_modes"]) == [2, 2, 2]
    # vals: [v, c, kept-row coefficients] per record
    assert np.array_equal(tape["vals"].view(np.uint64),
                          np.array([1, 1e16, -0.0, 1, 1, 2, 1, -1e16, 3, 4], np.float64).view(np.uint64))
    obj_skip, tape_skip = O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS,
                                                objective=False)
    assert obj_skip == 0.0
    for k in tape:
        assert np.array_equal(tape[k], tape_skip[k])
    assert O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS, with_tape=False)[0] == 1.0
    assert O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS, with_tape=False,
                                 objective=False)[0] == 0.0
    obj0, tape0 = O.fixed_variable_tape(FT_AT, [0] * 5, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS, objective=False)
    assert obj0 == 0.0 and len(tape0["types"]) == 0


def test_long_double_objective_differs_from_binary64():
    """The objective is not a binary64 sum: the same three products summed in
    double lose the 1 (a plain det_sum would give 0)."""
    assert (1e16 + 1.0) + -1e16 == 0.0
    assert O.fixed_variable_tape(FT_AT, FT_FIXED, FT_EXCLUDED, FT_VALUES, FT_ROWS, FT_COSTS, with_tape=False)[0] == 1.0


# ---- parallel-row hash closed form (cpp/tests/test_gpu_presolve_kernels.cpp:139-179) ----
FNV_OFFSET = 0xcbf29ce484222325
FNV_PRIME = 0x100000001b3
MASK64 = (1 << 64) - 1


def mix(h, x):
    return ((h ^ (x & MASK64)) * FNV_PRIME) & MASK64


def hash_fixture():
    lengths = [0, 1, 8, 9, 31, 32, 33, 1024, 1025, 4097]
    ptr, idx, vals = [0], [], []
    for row in range(2100):
        length = lengths[row] if row < 10 else 1
        for j in range(length):
            idx.append(j)
            vals.append(-0.5 if j % 2 else 2.0)
        ptr.append(len(idx))
    keep = np.ones(2100, np.uint8)
    keep[7] = 0
    return ptr, idx, vals, keep


def hash_expected(ptr, idx, vals, keep, rows):
    nnz = ptr[rows]
    use_xor = nnz > 32 * rows
    out = []
    for row in range(rows):
        first, last = ptr[row], ptr[row + 1]
        if not keep[row]:
            out.append(row)
            continue
        if first == last:
            out.append(mix(FNV_OFFSET, row))
            continue
        h = mix(FNV_OFFSET, last - first)
        agg = 0
        for p in range(first, last):
            bits = int(np.float64(vals[p] / vals[first]).view(np.uint64))
            if use_xor:
                agg ^= mix(mix(FNV_OFFSET, idx[p]), bits)
            else:
                h = mix(mix(h, idx[p]), bits)
        out.append(mix(h, agg) if use_xor else h)
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("rows", [10, 2100])
def test_parallel_row_hash_closed_form(rows):
    ptr, idx, vals, keep = hash_fixture()
    exp = hash_expected(ptr, idx, vals, keep, rows)
    got = O.parallel_row_hashes(rows, 4097, ptr[:rows + 1], idx[:ptr[rows]], vals[:ptr[rows]], keep[:rows])
    assert np.array_equal(got, exp)


# ---- stable-filter compaction oracle (cpp/tests/test_gpu_direct_port.cpp:71-147) ----
def compaction_fixture():
    m, n = 32768, 9000
    lengths = [0, 1, 8, 9, 31, 32, 33, 1024, 1025, 4097, 8193]
    col_map = np.full(n, -1, np.int64)
    expected_cols = [j for j in range(n) if j % 7]
    col_map[expected_cols] = np.arange(len(expected_cols))
    rp, ci, av = [0], [], []
    exp_rp, exp_ci, exp_av, row_map = [0], [], [], []
    for i in range(m):
        length = lengths[i] if i < 11 else 9
        before = len(exp_ci)
        for k in range(length):
            j = k // 2 if i < 11 else (i * 3 + k) % n
            v = -0.0 if k % 17 == 0 else (-0.1 if k % 3 == 0 else 0.3)
            ci.append(j)
            av.append(v)
            if col_map[j] >= 0:
                exp_ci.append(int(col_map[j]))
                exp_av.append(v)
        rp.append(len(ci))
        # empty-row removal counts stored entries, numerical zeros included
        if len(exp_ci) != before:
            row_map.append(i)
            exp_rp.append(len(exp_ci))
    upper = np.full(n, 10.0)
    upper[::7] = 0.0
    import gpupre as G
    lp = G.HostLP(m, n, np.asarray(rp, G.NP_OFFSET), np.asarray(ci, np.int32), np.asarray(av, np.float64),
                  np.zeros(n), np.zeros(n), upper, np.full(m, -1e9), np.full(m, 1e9))
    params = G.parity_params(max_iters=3, record_postsolve_tape=1, enable_startup=0, enable_singleton_rows=0,
                             enable_infeasible_and_redundant_rows=0, enable_bound_strengthening=0,
                             enable_parallel_rows=0, enable_empty_columns=0, enable_singleton_columns=0,
                             enable_doubleton_equations=0, enable_dual_fix=0, enable_parallel_columns=0)
    expect = dict(row_map=np.asarray(row_map, np.int32), col_map=np.asarray(expected_cols, np.int32),
                  row_ptr=np.asarray(exp_rp, G.NP_OFFSET), col_idx=np.asarray(exp_ci, np.int32),
                  vals=np.asarray(exp_av, np.float64))
    return lp, params, expect


def test_compaction_stable_filter_oracle():
    lp, params, e = compaction_fixture()
    r = O.presolve(lp, params)
    assert not r.has_infeasible and not r.has_unbounded
    assert np.array_equal(r.row_red2org, e["row_map"]) and np.array_equal(r.col_red2org, e["col_map"])
    assert (r.m1, r.n1, r.reduced.nnz) == (len(e["row_map"]), len(e["col_map"]), len(e["col_idx"]))
    assert np.array_equal(r.reduced.row_ptr, e["row_ptr"])
    assert np.array_equal(r.reduced.col_idx, e["col_idx"])
    assert np.array_equal(r.reduced.vals.view(np.uint64), e["vals"].view(np.uint64))
