"""LPs for the folding tests: the reference's folding test LPs
(cpp/tests/test_folding.cpp) and seeded LPs with a known symmetry -- k
identical copies of a random block, joined by linking rows that treat every
copy alike, then shuffled -- whose copies colour refinement merges."""
from __future__ import annotations

import numpy as np

import gpupre as G

INF = np.inf


def csr_lp(m, n, rows, cols, vals, c, l, u, AL, AU, obj_constant=0.0) -> G.HostLP:
    rows, cols, vals = np.asarray(rows, np.int64), np.asarray(cols, np.int32), np.asarray(vals, np.float64)
    order = np.lexsort((cols, rows))
    rp = np.searchsorted(rows[order], np.arange(m + 1)).astype(np.int32)
    return G.HostLP(m, n, rp, cols[order], vals[order], np.asarray(c, float), np.asarray(l, float),
                    np.asarray(u, float), np.asarray(AL, float), np.asarray(AU, float), float(obj_constant))


def symmetric_2x2(vals=(1.0, 1.0, 1.0, 1.0), c=(2.0, 2.0), AL=(4.0, 4.0), AU=(4.0, 4.0), obj_constant=0.0):
    """make_symmetric_2x2_lp (test_folding.cpp:68-87) with its variants."""
    return csr_lp(2, 2, [0, 0, 1, 1], [0, 1, 0, 1], vals, c, [0.0, 0.0], [10.0, 10.0], AL, AU, obj_constant)


def unequal_sizes():
    """test_unequal_color_sizes (test_folding.cpp:318-336): 2 x 3, all ones."""
    return csr_lp(2, 3, [0, 0, 0, 1, 1, 1], [0, 1, 2, 0, 1, 2], [1.0] * 6, [2.0] * 3, [0.0] * 3, [10.0] * 3,
                  [6.0, 6.0], [6.0, 6.0])


def diagonal_5x5():
    """test_rejection_preserves_passthrough_model (test_folding.cpp:411-437)."""
    return csr_lp(5, 5, range(5), range(5), [1, 2, 3, 4, 5], [0.0] * 5, [0.0] * 5, [INF] * 5, [0, 1, 2, 3, 4],
                  [INF] * 5)


def constant_columns(value, varied):
    """test_constant_reduction_matches_ordered_fallback (test_folding.cpp:439-471)
    as a whole folding run: no matrix entries, so every row and every column
    falls into one class (all values round to one float)."""
    n = 4097
    c = np.full(n, value)
    lower = np.full(n, value)
    upper = np.full(n, INF)
    if varied:
        with np.errstate(over="ignore"):  # the largest double steps to inf, as in the reference test
            c[n // 2] = np.nextafter(value, INF)
            lower[n // 3] = np.nextafter(value, -INF)
    AL = np.array([0.1, 0.1, np.nextafter(0.1, INF)])
    return csr_lp(3, n, [], [], [], c, lower, upper, AL, [INF] * 3)


def symmetric_lp(seed, copies=4, p=6, q=9, links=3, integer=False):
    """`copies` identical copies of a random p x q block (rows two-sided or
    equalities, bounded columns) plus `links` linking rows, each summing one
    block column over every copy, rows and columns shuffled.  Returns the LP
    and the number of copies."""
    rng = np.random.default_rng(seed)
    nb = int(p * q * 0.4)
    flat = rng.choice(p * q, size=nb, replace=False)
    bi, bj = flat // q, flat % q
    bv = rng.integers(1, 6, nb) * rng.choice([-1, 1], nb) if integer else rng.uniform(0.5, 3.0, nb)
    bc = rng.integers(-5, 6, q).astype(float) if integer else rng.normal(size=q)
    bl = np.zeros(q)
    bu = rng.integers(2, 9, q).astype(float)
    x0 = bl + (bu - bl) * (rng.integers(0, 5, q) / 4.0 if integer else rng.uniform(0.2, 0.8, q))
    act = np.zeros(p)
    np.add.at(act, bi, bv * x0[bj])
    bAL = act - rng.integers(0, 3, p)
    bAU = act + rng.integers(0, 3, p)
    rows, cols, vals = [], [], []
    for k in range(copies):
        rows += list(bi + k * p)
        cols += list(bj + k * q)
        vals += list(bv)
    m, n = copies * p + links, copies * q
    link_cols = rng.choice(q, size=links, replace=False)
    AL = list(np.tile(bAL, copies))
    AU = list(np.tile(bAU, copies))
    for t, j in enumerate(link_cols):
        rows += [copies * p + t] * copies
        cols += [j + k * q for k in range(copies)]
        vals += [1.0] * copies
        s = copies * x0[j]
        AL.append(s - 1.0)
        AU.append(s + 1.0)
    rows, cols = np.asarray(rows), np.asarray(cols)
    rperm, cperm = rng.permutation(m), rng.permutation(n)  # new index of old row / column
    c = np.empty(n)
    l = np.empty(n)
    u = np.empty(n)
    c[cperm] = np.tile(bc, copies)
    l[cperm] = np.tile(bl, copies)
    u[cperm] = np.tile(bu, copies)
    AL2, AU2 = np.empty(m), np.empty(m)
    AL2[rperm] = AL
    AU2[rperm] = AU
    return csr_lp(m, n, rperm[rows], cperm[cols], vals, c, l, u, AL2, AU2, obj_constant=1.5), copies
