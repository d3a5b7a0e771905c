"""Learned pass scheduler on the CPU: the restated network and its gradient,
the feature vector of the engine against the restatement, and the gated
schedule of the restatement.

Pins:
* forward: a float64 NumPy evaluation of the same network (independent
  matrix code) agrees to FP32 rounding;
* gradient: central finite differences of the float64 loss agree with the
  restatement's analytic FP32 gradient (a dropped term, a wrong index or a
  missing ReLU mask fails this);
* Adam: the closed form of the first step (|update| = lr for every nonzero
  gradient);
* schedule: all groups on reproduces the plain schedule bit for bit; all off
  defers every reduction to the plain fixpoint sweep with the same result.
"""
import os

import numpy as np
import pytest

import gpupre as G
import learned as L
import oracle as O
from lp_check import primal_residual, relative_gap, solve

IN, H1, H2, OUT = G.GP_MLP_IN, G.GP_MLP_H1, G.GP_MLP_H2, G.GP_MLP_OUT


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    k = 0
    out = []
    for fi, fo in ((IN, H1), (H1, H2), (H2, OUT)):
        W = P[k:k + fo * fi].reshape(fo, fi)
        k += fo * fi
        b = P[k:k + fo]
        k += fo
        out += [W, b]
    assert k == G.GP_MLP_PARAMS
    return out


def np_forward(P, X):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


def np_loss(P, X, Y, M):
    o = np_forward(P, X)
    ell = np.maximum(o, 0.0) - o * Y + np.log1p(np.exp(-np.abs(o)))
    return float((ell * M).sum() / max(1.0, M.sum()))


def batch(seed, B=40):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, size=(B, IN)).astype(np.float32)
    Y = (rng.uniform(size=(B, OUT)) < 0.3).astype(np.float32)
    M = (rng.uniform(size=(B, OUT)) < 0.6).astype(np.float32)
    P = L.init_params(seed)
    P[G.GP_MLP_PARAMS - OUT:] = rng.normal(0.0, 0.5, OUT).astype(np.float32)  # nonzero output bias
    return P, X, Y, M


def test_forward_matches_float64():
    P, X, _, _ = batch(1)
    got = O.mlp_forward(P, X).astype(np.float64)
    ref = np_forward(P, X.astype(np.float64))
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)


def test_loss_and_gradient_match_finite_differences():
    P, X, Y, M = batch(2)
    loss, grad = O.mlp_loss_grad(P, X, Y, M)
    assert abs(loss - np_loss(P, X.astype(np.float64), Y, M)) <= 1e-5 * max(1.0, abs(loss))
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(G.GP_MLP_PARAMS, 60, replace=False),
                          np.arange(G.GP_MLP_PARAMS - OUT, G.GP_MLP_PARAMS)])  # every output bias
    P64 = P.astype(np.float64)
    X64 = X.astype(np.float64)
    h = 1e-5
    for i in idx:
        Pp, Pm = P64.copy(), P64.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (np_loss(Pp, X64, Y, M) - np_loss(Pm, X64, Y, M)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-4 * max(1e-2, abs(fd)), (i, fd, grad[i])


def test_masked_entries_do_not_contribute():
    P, X, Y, M = batch(4)
    Y2 = Y.copy()
    Y2[M == 0] = 1.0 - Y2[M == 0]
    l1, g1 = O.mlp_loss_grad(P, X, Y, M)
    l2, g2 = O.mlp_loss_grad(P, X, Y2, M)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_adam_first_step_closed_form():
    P, X, Y, M = batch(5)
    _, g = O.mlp_loss_grad(P, X, Y, M)
    w = P.copy()
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    O.mlp_adam_step(w, g, m, v, 1, 1e-3)
    step = (P - w).astype(np.float64)
    nz = np.abs(g) > 1e-4  # eps = 1e-8 changes these steps by < 1e-4 relative
    assert np.allclose(np.abs(step[nz]), 1e-3, rtol=1e-3)
    assert np.all(np.sign(step[nz]) == np.sign(g[nz]))


def test_rule_groups_engine_and_restatement():
    lib, olib = G.load(), O.load()
    for r in range(11):
        assert lib.gp_rule_group(r) == olib.or_rule_group(r) == L.GROUP_OF_RULE[r]


def lps():
    return [("qap6", G.generate_qap(6, seed=3)),
            ("gen4k", G.generate(4000, 6000, seed=5, degree_mode=0, mean_degree=6.0))]


@pytest.mark.parametrize("name,lp", lps(), ids=[x[0] for x in lps()])
def test_engine_features_match_restatement(name, lp):
    r = O.presolve(lp, G.parity_params(collect_trace=1))
    X, _, _ = L.round_samples(r.trace, lp.m, lp.n)
    nnz0 = r.trace[0].nnz_before
    for t, tr in enumerate(r.trace):
        f = O.schedule_features(r.trace[:t], t, tr.phase, tr.nnz_before, nnz0, lp.m, lp.n,
                                tr.bound_only_rounds, tr.bound_only_limit)
        assert np.array_equal(f.view(np.uint32), X[t].view(np.uint32)), (name, t)


def same_result(a, b):
    assert (a.m1, a.n1, a.reduced.nnz) == (b.m1, b.n1, b.reduced.nnz)
    assert np.array_equal(a.row_red2org, b.row_red2org) and np.array_equal(a.col_red2org, b.col_red2org)
    assert np.array_equal(a.types, b.types) and np.array_equal(a.indices, b.indices)
    assert np.array_equal(a.vals.view(np.uint64), b.vals.view(np.uint64))
    assert np.array_equal(a.reduced.vals.view(np.uint64), b.reduced.vals.view(np.uint64))


@pytest.mark.parametrize("name,lp", lps(), ids=[x[0] for x in lps()])
def test_gate_all_on_is_plain_and_all_off_defers_to_fixpoint_sweep(name, lp):
    plain = O.presolve(lp, G.parity_params(collect_trace=1))
    P = L.init_params(0)
    on = L.make_blob(P, np.full(OUT, -np.inf, np.float32))
    r_on = O.presolve(lp, L.learned_params(on, collect_trace=1))
    same_result(plain, r_on)
    assert all(t.skipped_mask == 0 for t in r_on.trace)
    off = L.make_blob(P, np.full(OUT, np.inf, np.float32))
    r_off = O.presolve(lp, L.learned_params(off, collect_trace=1))
    gated = [t for t in r_off.trace if not t.final_sweep]
    assert [t.phase for t in gated] == [0, 1, 2, 3] and all(t.fired_mask == 0 for t in gated)
    same_result(plain, r_off)
    assert L.rule_launches(r_off.trace) == L.rule_launches(plain.trace)


def test_learned_weights_file():
    if not os.path.exists(L.WEIGHTS_PATH):
        pytest.skip("no trained scheduler weights in the package")
    blob = L.load_blob()
    assert blob.size == G.GP_MLP_BLOB and np.all(np.isfinite(blob[:G.GP_MLP_PARAMS]))


@pytest.mark.parametrize("name,lp", lps(), ids=[x[0] for x in lps()])
def test_learned_schedule_postsolve_gap(name, lp):
    if not os.path.exists(L.WEIGHTS_PATH):
        pytest.skip("no trained scheduler weights in the package")
    blob = L.load_blob()
    st0, obj0, *_ = solve(lp)
    assert st0 == 0
    res = O.presolve(lp, L.learned_params(blob))
    assert not res.has_infeasible and not res.has_unbounded
    st, obj_r, xr, yr, zr = solve(res.reduced)
    assert st == 0
    x, _, _ = O.postsolve(res, lp, xr, yr, zr)
    assert relative_gap(float(lp.c @ x) + lp.obj_constant, obj0) <= 1e-6
    assert primal_residual(lp, x) <= 1e-6
