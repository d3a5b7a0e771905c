"""The Adams-Johnson QAP LP generator (scripts/generate_qap20.jl:25-109).

Pinned without running the Julia script:
* the row / column / nnz counts of its closed forms (the script prints the
  first two, generate_qap20.jl:116);
* every permutation of a small QAP gives a feasible point of the LP
  (x = permutation matrix, y_ijkl = x_ij x_kl), whose objective is the QAP
  cost of that permutation -- which fixes the constraint structure and the
  cost layout y[i,j,k,l] -> flow[i,k] * distance[j,l];
* the LP optimum (HiGHS) is a lower bound of the best permutation, and the
  planted point 1/n, 1/n^2 is feasible;
* the presolve of the CPU restatement keeps the LP optimum (postsolve gap).
"""
import itertools

import numpy as np
import pytest

import gpupre as G
import oracle as O
from lp_check import csr, primal_residual, relative_gap, solve


def perm_point(n, perm):
    x = np.zeros(n * n + n ** 4)
    for i in range(n):
        x[i + n * perm[i]] = 1.0
    for i in range(n):
        for k in range(n):
            j, l = perm[i], perm[k]
            x[n * n + i + n * j + n * n * k + n ** 3 * l] = 1.0
    return x


@pytest.mark.parametrize("n", [2, 3, 4, 5, 12])
def test_qap_shape(n):
    lp = G.generate_qap(n, seed=n)
    assert lp.m == 2 * n ** 4 + 2 * n ** 3 + n * n + 2 * n
    assert lp.n == n ** 4 + n * n
    assert lp.nnz == n * n * (2 * n + 3) + 3 * n ** 4 + 2 * (n ** 4 - n * n)
    A = csr(lp)
    assert A.has_sorted_indices
    assert np.all(np.isinf(lp.l)) and np.all(np.isinf(lp.u))
    assert primal_residual(lp, lp.x0) <= 1e-12


def test_qap_permutations_feasible_and_costed():
    n = 4
    lp = G.generate_qap(n, seed=7)
    A = csr(lp)
    costs = []
    for perm in itertools.permutations(range(n)):
        x = perm_point(n, perm)
        ax = A @ x
        assert np.all(ax >= lp.AL - 1e-12) and np.all(ax <= lp.AU + 1e-12), perm
        costs.append(float(lp.c @ x))
    # the y costs of a permutation point are flow[i,k] * dist[perm i, perm k]:
    # recover flow (x) dist from c and check the rank-one product structure
    cy = lp.c[n * n:].reshape(n, n, n, n, order="F")  # cy[i, j, k, l]
    fd = cy.transpose(0, 2, 1, 3).reshape(n * n, n * n)  # rows (i, k), cols (j, l)
    assert np.linalg.matrix_rank(fd) <= 1
    assert np.all(cy[np.arange(n), :, np.arange(n), :] == 0)  # zero flow diagonal
    assert np.all(cy[:, np.arange(n), :, np.arange(n)] == 0)  # zero distance diagonal
    st, obj, *_ = solve(lp)
    assert st == 0
    assert obj <= min(costs) + 1e-6 * max(1.0, abs(min(costs)))
    assert obj >= -1e-9


@pytest.mark.parametrize("n", [3, 5])
def test_qap_presolve_keeps_optimum(n):
    lp = G.generate_qap(n, seed=11)
    st0, obj0, *_ = solve(lp)
    assert st0 == 0
    res = O.presolve(lp, G.parity_params())
    assert not res.has_infeasible and not res.has_unbounded
    assert res.m1 < lp.m
    st, obj_r, xr, yr, zr = solve(res.reduced)
    assert st == 0
    x, _, _ = O.postsolve(res, lp, xr, yr, zr)
    assert relative_gap(float(lp.c @ x) + lp.obj_constant, obj0) <= 1e-6
    assert primal_residual(lp, x) <= 1e-6
