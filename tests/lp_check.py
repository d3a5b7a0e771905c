"""End-to-end checks of presolve + postsolve against an LP solver.

The reduced LP is solved with SciPy's HiGHS, its primal/dual solution is
mapped back through postsolve, and the recovered point is checked on the
original LP: objective gap against HiGHS on the original model, primal and
dual feasibility.  Sign convention c - A'y - z = 0 (the engine's)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

import gpupre as G


def csr(lp: G.HostLP):
    return sp.csr_matrix((lp.vals, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))


def solve(lp: G.HostLP):
    """HiGHS on min c'x + k s.t. AL <= Ax <= AU, l <= x <= u.  Returns
    (status, objective, x, y, z) with y, z in the engine's convention."""
    A = csr(lp)
    AL, AU = np.asarray(lp.AL), np.asarray(lp.AU)
    eq = np.isfinite(AL) & np.isfinite(AU) & (AL == AU)
    up = np.isfinite(AU) & ~eq
    lo = np.isfinite(AL) & ~eq
    blocks, rhs = [], []
    if up.any():
        blocks.append(A[up])
        rhs.append(AU[up])
    if lo.any():
        blocks.append(-A[lo])
        rhs.append(-AL[lo])
    A_ub = sp.vstack(blocks).tocsr() if blocks else None
    b_ub = np.concatenate(rhs) if rhs else None
    A_eq = A[eq] if eq.any() else None
    b_eq = AU[eq] if eq.any() else None
    bounds = np.column_stack([np.where(np.isfinite(lp.l), lp.l, -np.inf), np.where(np.isfinite(lp.u), lp.u, np.inf)])
    res = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, bounds=bounds, method="highs",
                  options={"primal_feasibility_tolerance": 1e-9, "dual_feasibility_tolerance": 1e-9})
    if res.status != 0:
        return res.status, None, None, None, None
    y = np.zeros(lp.m)
    k = 0
    mu = res.ineqlin.marginals if A_ub is not None else np.zeros(0)
    nup = int(up.sum())
    y[up] += mu[:nup]
    y[lo] -= mu[nup:]
    if A_eq is not None:
        y[eq] = res.eqlin.marginals
    z = res.lower.marginals + res.upper.marginals
    obj = float(res.fun) + float(getattr(lp, "obj_constant", 0.0))
    return 0, obj, res.x, y, z


def primal_residual(lp: G.HostLP, x):
    ax = csr(lp) @ x
    r = np.maximum(0.0, np.maximum(np.asarray(lp.AL) - ax, ax - np.asarray(lp.AU)))
    b = np.maximum(0.0, np.maximum(np.asarray(lp.l) - x, x - np.asarray(lp.u)))
    r = r[np.isfinite(r)]
    b = b[np.isfinite(b)]
    return float(max(r.max(initial=0.0), b.max(initial=0.0)))


def dual_residual(lp: G.HostLP, y, z):
    return float(np.abs(np.asarray(lp.c) - csr(lp).T @ y - z).max(initial=0.0))


def relative_gap(a, b):
    return abs(a - b) / max(1.0, abs(a), abs(b))
