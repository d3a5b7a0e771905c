"""LPs for the structural Startup rules: the reference's test LPs
(cpp/tests/test_rule_column_domination.cpp, test_rule_projected_auxiliaries.cpp)
and seeded LPs with planted structure -- unit set covering with dominated
columns, zero-cost two-row columns and McCormick products among decoys that
each break one condition of the certificates -- rows and columns shuffled."""
from __future__ import annotations

import numpy as np

import gpupre as G
from fold_cases import csr_lp

INF = np.inf


def covering_lp(variant="valid"):
    """make_covering_lp (test_rule_column_domination.cpp:68-106): 3 x 6 unit
    covering, costs (1, 1, 1, 3, 3, 4); columns 3 and 5 exceed their degree."""
    rows = [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2]
    cols = [0, 3, 4, 5, 1, 3, 4, 5, 2, 4, 5]
    vals = [1.0] * 11
    c = [1.0, 1.0, 1.0, 3.0, 3.0, 4.0]
    AU = [INF] * 3
    u = [1.0] * 6
    if variant == "nonunit":
        vals[1] = 0.5  # A(0, 3); the transpose carries the same entry
    elif variant == "noncovering":
        AU[0] = 1.0
    elif variant == "missing_witness":
        c[2] = 2.0
    elif variant == "nonbinary":
        u[5] = 2.0
    return csr_lp(3, 6, rows, cols, vals, c, [0.0] * 6, u, [1.0] * 3, AU)


def pair_block_lp(second_rhs=4.0, recourse_cost=0.0):
    """make_pair_block_lp (test_rule_projected_auxiliaries.cpp:121-145): a
    McCormick product x4 = x0 x1 and a zero-cost two-row column x5."""
    rp = [0, 3, 5, 7, 11, 15]
    ci = [0, 1, 4, 0, 4, 1, 4, 2, 3, 4, 5, 2, 3, 4, 5]
    av = [-1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 2.0, -2.0, -1.0, 1.0, 2.0, 2.0]
    rows = np.repeat(np.arange(5), np.diff(rp))
    return csr_lp(5, 6, rows, ci, av, [-1.0, -2.0, 0.0, 0.0, 0.0, recourse_cost], [0.0] * 6, [1.0] * 6,
                  [-1.0, -INF, -INF, -INF, -INF], [INF, 0.0, 0.0, 2.0, second_rhs])


def mccormick_lp(corrupt_upper_row=False, nonunit_endpoint=False):
    """make_mccormick_lp (test_rule_projected_auxiliaries.cpp:147-166)."""
    rows = [0, 0, 0, 1, 1, 2, 2]
    ci = [0, 1, 2, 0, 2, 1, 2]
    av = [-1.0, -1.0, 1.0, -0.5 if corrupt_upper_row else -1.0, 1.0, -1.0, 1.0]
    u = [1.0, 2.0 if nonunit_endpoint else 1.0, 1.0]
    return csr_lp(3, 3, rows, ci, av, [0.0] * 3, [0.0] * 3, u, [-1.0, -INF, -INF], [INF, 0.0, 0.0])


def pending_lp():
    """test_pending_masks_and_gpu_tape (test_rule_projected_auxiliaries.cpp:339-398):
    the pair block with an extra first row and an extra column 6 in every row,
    to be presolved against a plan that already deletes row 0 and column 6."""
    base = pair_block_lp()
    rows, cols, vals = [0, 0, 0, 0], [0, 4, 5, 6], [1.0, 1.0, 1.0, 1.0]
    for i in range(5):
        for p in range(base.row_ptr[i], base.row_ptr[i + 1]):
            rows.append(i + 1)
            cols.append(int(base.col_idx[p]))
            vals.append(float(base.vals[p]))
        rows.append(i + 1)
        cols.append(6)
        vals.append(123.0)
    return csr_lp(6, 7, rows, cols, vals, [-1.0, -2.0, 0.0, 0.0, 0.0, 0.0, 3.0], [0.0] * 7,
                  [1.0, 1.0, 1.0, 1.0, 1.0, 0.25, 1.0], [-INF, -1.0, -INF, -INF, -INF, -INF],
                  [INF, INF, 0.0, 0.0, 2.0, 4.0])


def startup_params(**kw):
    """The reference tests' settings: presolve alone (no folding), one
    iteration with every general rule off, so the Startup rules alone act."""
    base = dict(max_iters=1, enable_folding=0, enable_piecewise_linear_reformulation_variable_pair=0,
                enable_piecewise_linear_reformulation_absolute_value=0, enable_infeasible_and_fixed_variables=0,
                enable_empty_rows=0, enable_singleton_rows=0, enable_infeasible_and_redundant_rows=0,
                enable_bound_strengthening=0, enable_parallel_rows=0, enable_empty_columns=0,
                enable_singleton_columns=0, enable_doubleton_equations=0, enable_dual_fix=0,
                enable_parallel_columns=0)
    base.update(kw)
    return G.default_params(**base)


def covering_params(**kw):
    """test_rule_column_domination.cpp:108-133"""
    base = dict(column_domination_min_fixed_cols=1, column_domination_min_candidate_ratio=0.20,
                column_domination_probe_rows=3, column_domination_probe_cols=6, column_domination_probe_nnz=11)
    base.update(kw)
    return startup_params(**base)


def projected_params(**kw):
    """projected_only_params (test_rule_projected_auxiliaries.cpp:168-198)"""
    base = dict(enable_column_domination=0, two_row_variable_elimination_min_candidates=1,
                two_row_variable_elimination_min_candidate_ratio=0.0, zero_cost_product_variables_min_candidates=1,
                zero_cost_product_variables_min_candidate_ratio=0.0)
    base.update(kw)
    return startup_params(**base)


# ------------------------------------------------------------------ seeded LPs
def planted_covering(seed, m=400, n=6000, cost_max=9, density=0.004):
    """Unit set covering: every row >= 1 over binary columns with positive
    integer costs and a cost-one column in every row; columns whose cost
    exceeds their degree are the dominated ones."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for j in range(n):
        deg = max(1, rng.binomial(m, density))
        for i in rng.choice(m, size=min(deg, m), replace=False):
            rows.append(int(i))
            cols.append(j)
    c = rng.integers(1, cost_max + 1, n).astype(float)
    # a cost-one witness in every row
    for i in range(m):
        j = int(rng.integers(0, n))
        c[j] = 1.0
        rows.append(i)
        cols.append(j)
    pairs = np.unique(np.stack([rows, cols], 1), axis=0)
    return csr_lp(m, n, pairs[:, 0], pairs[:, 1], np.ones(len(pairs)), c, np.zeros(n), np.ones(n), np.ones(m),
                  np.full(m, INF))


def planted_projections(seed, blocks=40, extra_rows=30, decoy_every=4):
    """Binary columns x linked by random covering rows, with planted
    McCormick products z = x_a x_b (zero cost) and zero-cost two-row columns
    w (the pair-block certificate over random x's, rows scaled by random
    factors), plus decoys: a two-row column whose second side is tight
    (rhs 3), one with a cost, a product over a non-binary endpoint, and a
    product chained onto another product.  Returns the LP."""
    rng = np.random.default_rng(seed)
    nx = 6 * blocks + 8
    ents = []  # (row, col, val)
    AL, AU = [], []
    c, l, u = list(rng.integers(-5, 6, nx).astype(float)), [0.0] * nx, [1.0] * nx
    for k in (3, 7):
        u[k] = 2.0  # non-binary x's for product decoys
    ncol = nx

    def new_col(cost=0.0, lo=0.0, hi=1.0):
        nonlocal ncol
        c.append(cost)
        l.append(lo)
        u.append(hi)
        ncol += 1
        return ncol - 1

    def add_row(entries, lo, hi):
        r = len(AL)
        for j, a in entries:
            ents.append((r, j, a))
        AL.append(lo)
        AU.append(hi)

    products = []
    for b in range(blocks):
        decoy = (b % decoy_every) == decoy_every - 1
        kind = int(rng.integers(0, 4)) if decoy else -1
        # McCormick product
        xa, xb = (int(v) for v in rng.choice(nx, 2, replace=False))
        if kind == 2:
            xb = 3 if xa != 3 else 7
        if kind == 3 and products:
            xb = products[-1]  # chained onto an earlier product (both cols are candidates)
        z = new_col()
        products.append(z)
        s = float(rng.choice([1.0, -1.0, 2.0, -0.5, 4.0]))
        add_row([(z, s), (xa, -s), (xb, -s)], -s if s > 0 else -INF, INF if s > 0 else -s)
        for e in (xa, xb):
            t = float(rng.choice([1.0, -1.0, 3.0]))
            add_row([(z, t), (e, -t)], -INF if t > 0 else 0.0, 0.0 if t > 0 else INF)
        # zero-cost two-row column (pair-block certificate)
        x2, x3, x4 = (int(v) for v in rng.choice(nx, 3, replace=False))
        w = new_col(cost=0.25 if kind == 1 else 0.0)
        rhs = 3.0 if kind == 0 else 4.0
        s1, s2 = float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.5, 3.0))
        add_row([(x2, s1), (x3, -s1), (x4, 2 * s1), (w, -2 * s1)], -INF, 2 * s1)
        add_row([(x2, -s2), (x3, s2), (x4, 2 * s2), (w, 2 * s2)], -INF, rhs * s2)
    for _ in range(extra_rows):
        k = int(rng.integers(2, 6))
        add_row([(int(j), 1.0) for j in rng.choice(nx, k, replace=False)], 1.0, INF)
    m = len(AL)
    ents.sort()
    rows = [e[0] for e in ents]
    cols = [e[1] for e in ents]
    vals = [e[2] for e in ents]
    lp = csr_lp(m, ncol, rows, cols, vals, c, l, u, AL, AU)
    return shuffle_lp(lp, rng)


def shuffle_lp(lp, rng):
    m, n = lp.m, lp.n
    rperm, cperm = rng.permutation(m), rng.permutation(n)
    rows = np.repeat(np.arange(m), np.diff(lp.row_ptr))
    c, l, u = np.empty(n), np.empty(n), np.empty(n)
    c[cperm], l[cperm], u[cperm] = lp.c, lp.l, lp.u
    AL, AU = np.empty(m), np.empty(m)
    AL[rperm], AU[rperm] = lp.AL, lp.AU
    return csr_lp(m, n, rperm[rows], cperm[lp.col_idx], lp.vals, c, l, u, AL, AU, lp.obj_constant)
