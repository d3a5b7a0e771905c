"""Full-pipeline parity (GPU against the CPU restatement, bit for bit) on LPs
shaped to drive the less common branches of the engine's kernels:

* mass fixing -- most columns fixed at once, so the marked columns of the
  row shifts hold more than a quarter of the entries and the rows-touching
  lister sweeps A instead of walking the marked columns in AT;
* long columns -- columns of a few thousand up to all rows, so the
  entry-parallel segment reductions over AT (column max, dual-fix side
  flags, doubleton column counts) see segments crossing several tiles, and
  the compaction tiles hold single long lines.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gpupre as G
from test_gpu_parity import compare_pipeline

pytestmark = pytest.mark.gpu


def lp_from_columns(rng, m, n, col_rows, x0_scale=1.0):
    """An LP with the given column supports, a planted point x0 inside the
    bounds and row sides around A x0 (a third of them equalities); free
    columns cost nothing, so the LP is bounded.  Returns (lp, x0)."""
    r = np.concatenate(col_rows)
    c = np.concatenate([np.full(len(rows), j) for j, rows in enumerate(col_rows)])
    vals = rng.integers(1, 5, size=r.size) * rng.choice([-1.0, 1.0], size=r.size)
    A = sp.csr_matrix((vals, (r, c)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    x0 = np.round(rng.uniform(-x0_scale, x0_scale, size=n))
    l = x0 - rng.integers(0, 3, size=n)
    u = x0 + rng.integers(0, 3, size=n)
    free = rng.random(n) < 0.1
    l[free], u[free] = -np.inf, np.inf
    ax = A @ x0
    AL = ax - rng.integers(0, 3, size=m)
    AU = ax + rng.integers(0, 3, size=m)
    eq = rng.random(m) < 1 / 3
    AL[eq] = AU[eq] = ax[eq]
    AL[rng.random(m) < 0.1] = -np.inf
    AU[rng.random(m) < 0.1] = np.inf
    cost = rng.integers(-3, 4, size=n).astype(np.float64)
    cost[free] = 0.0
    lp = G.HostLP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64), cost,
                  l, u, AL, AU)
    return lp, x0


def short_columns(rng, m, n, lo=1, hi=12):
    return [rng.choice(m, size=int(rng.integers(lo, hi)), replace=False) for _ in range(n)]


@pytest.mark.parametrize("seed", [0, 1])
def test_pipeline_parity_mass_fixing(seed):
    rng = np.random.default_rng(300 + seed)
    m, n = 6000, 9000
    lp, x0 = lp_from_columns(rng, m, n, short_columns(rng, m, n))
    fix = rng.random(n) < 0.6  # fixed at the planted point: the LP stays feasible
    l, u = lp.l.copy(), lp.u.copy()
    l[fix] = u[fix] = x0[fix]
    lp = G.HostLP(lp.m, lp.n, lp.row_ptr, lp.col_idx, lp.vals, lp.c, l, u, lp.AL, lp.AU)
    s, ores, rec = compare_pipeline(lp, G.parity_params(), f"mass fixing seed={seed}")
    s.free()


@pytest.mark.parametrize("seed", [0, 1])
def test_pipeline_parity_long_columns(seed):
    rng = np.random.default_rng(400 + seed)
    m, n = 20000, 3000
    cols = short_columns(rng, m, n)
    cols[0] = np.arange(m)                                    # a column in every row
    for j in range(1, 12):                                    # columns over 2-4 segment tiles and beyond
        cols[j] = rng.choice(m, size=int(rng.integers(2000, 9000)), replace=False)
    lp, _ = lp_from_columns(rng, m, n, cols)
    s, ores, rec = compare_pipeline(lp, G.parity_params(), f"long columns seed={seed}")
    s.free()
