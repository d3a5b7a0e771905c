"""The synthetic generator plants a point x0 that satisfies every row and
every bound of the LP it writes, for the integer recipe of the 5M x 8M
workload and for the FP64 recipe (checked at 1/40 scale, same code path).
An infeasibility verdict on these LPs is therefore a tolerance artefact of
the presolver that reports it, not a property of the LP."""
import numpy as np
import pytest

import gpupre as G


def row_activity(lp, x):
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    return np.bincount(rows, weights=lp.vals * x[lp.col_idx], minlength=lp.m)


@pytest.mark.parametrize("workload,exact", [("large", True), ("large_fp64", False), ("medium", False)])
def test_planted_point_satisfies_rows_and_bounds(workload, exact):
    w = dict(G.WORKLOADS[workload])
    m, n = w.pop("m"), w.pop("n")
    scale = 40 if m >= 1_000_000 else 1
    lp = G.generate(m // scale, n // scale, 0, **w)
    x = lp.x0
    assert x is not None and x.size == lp.n
    assert np.all(x >= lp.l) and np.all(x <= lp.u)
    act = row_activity(lp, x)
    if exact:
        # integer data: activities are exact integers in binary64
        assert np.all(act == np.round(act))
        assert np.all(act >= lp.AL) and np.all(act <= lp.AU)
    else:
        tol = 1e-9 * (1.0 + np.abs(act))
        assert np.all(act >= lp.AL - tol) and np.all(act <= lp.AU + tol)
