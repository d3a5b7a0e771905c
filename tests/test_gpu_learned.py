"""Learned pass scheduler on the GPU: the HIP MLP kernels against the CPU
restatement, the gated presolve bit for bit against the restatement's gated
schedule, and what the trained network gains over the plain schedule.

Tolerances: the forward pass is bit-exact (same FP32 summation order, no FMA
contraction on either side); loss and gradient go through expf/log1pf, whose
last bits differ between the device library and libm, so they are compared
at 1e-5 relative to the largest gradient entry."""
import os
import time

import numpy as np
import pytest

import gpupre as G
import learned as L
import oracle as O
from lp_check import primal_residual, relative_gap, solve

pytestmark = pytest.mark.gpu

OUT = G.GP_MLP_OUT


def rand_batch(seed, B):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, size=(B, G.GP_MLP_IN)).astype(np.float32)
    Y = (rng.uniform(size=(B, OUT)) < 0.3).astype(np.float32)
    M = (rng.uniform(size=(B, OUT)) < 0.6).astype(np.float32)
    P = L.init_params(seed)
    P[G.GP_MLP_PARAMS - OUT:] = rng.normal(0.0, 0.5, OUT).astype(np.float32)
    return P, X, Y, M


@pytest.mark.parametrize("B", [1, 7, 1000])
def test_mlp_forward_bit_exact(B):
    P, X, _, _ = rand_batch(B, B)
    tr = L.GpuTrainer(P)
    got = tr.logits(X)
    tr.free()
    ref = O.mlp_forward(P, X)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("B", [3, 512])
def test_mlp_loss_grad_and_adam_match_restatement(B):
    P, X, Y, M = rand_batch(10 + B, B)
    tr = L.GpuTrainer(P)
    tr.set_data(X, Y, M)
    lib = G.load()
    G.check(lib.gp_mlp_loss_grad(tr.P.ptr, tr.data[0].ptr, tr.data[1].ptr, tr.data[2].ptr, B, tr.loss.ptr,
                                 tr.grad.ptr))
    g_gpu = G.to_host(tr.grad.ptr, G.GP_MLP_PARAMS, np.float32)
    loss_gpu = float(G.to_host(tr.loss.ptr, 1, np.float32)[0])
    loss_ref, g_ref = O.mlp_loss_grad(P, X, Y, M)
    scale = float(np.abs(g_ref).max())
    assert np.abs(g_gpu.astype(np.float64) - g_ref).max() <= 1e-5 * scale
    assert abs(loss_gpu - loss_ref) <= 1e-5 * max(1.0, abs(loss_ref))
    # one Adam step from the same gradient
    G.check(lib.gp_mlp_adam_step(tr.P.ptr, tr.grad.ptr, tr.m.ptr, tr.v.ptr, 1, 1e-3, 0.9, 0.999, 1e-8))
    w_gpu = tr.params()
    tr.free()
    w = P.copy()
    O.mlp_adam_step(w, g_gpu, np.zeros_like(P), np.zeros_like(P), 1, 1e-3)
    assert np.abs(w_gpu.astype(np.float64) - w).max() <= 1e-6


def test_training_reduces_loss():
    """A few hundred GPU Adam steps fit a separable labelling."""
    rng = np.random.default_rng(0)
    X = rng.uniform(0.0, 1.0, size=(2048, G.GP_MLP_IN)).astype(np.float32)
    Y = np.stack([(X[:, g] > 0.5) for g in range(OUT)], axis=1).astype(np.float32)
    M = np.ones_like(Y)
    tr = L.GpuTrainer(L.init_params(1))
    tr.set_data(X, Y, M)
    first = tr.step(0, 256, 3e-3)
    for ep in range(40):
        for s in range(0, 2048, 256):
            last = tr.step(s, 256, 3e-3)
    acc = ((tr.logits(X) >= 0) == (Y > 0)).mean()
    tr.free()
    assert last < 0.25 * first and acc > 0.95


def weights():
    if not os.path.exists(L.WEIGHTS_PATH):
        pytest.skip("no trained scheduler weights in the package")
    return L.load_blob()


def lps():
    return [("qap8", G.generate_qap(8, seed=101)),
            ("gen20k", G.generate(20000, 30000, seed=103, degree_mode=0, mean_degree=6.0))]


def gpu_run(lp, params):
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            rec = s.record()
            return rec, s.reduced(), s.trace(), (s.reduced_rows, s.reduced_cols, s.reduced_nnz)
        finally:
            s.free()


@pytest.mark.parametrize("name,lp", lps(), ids=[x[0] for x in lps()])
def test_learned_presolve_matches_restatement(name, lp):
    blob = weights()
    params = L.learned_params(blob, collect_trace=1)
    rec, red, trace, dims = gpu_run(lp, params)
    ref = O.presolve(lp, params)
    assert dims == (ref.m1, ref.n1, ref.reduced.nnz)
    assert [(t.phase, t.fired_mask, t.skipped_mask, t.final_sweep) for t in trace] == \
           [(t.phase, t.fired_mask, t.skipped_mask, t.final_sweep) for t in ref.trace]
    assert np.array_equal(rec.row_red2org, ref.row_red2org) and np.array_equal(rec.col_red2org, ref.col_red2org)
    assert np.array_equal(rec.types, ref.types) and np.array_equal(rec.indices, ref.indices)
    assert np.array_equal(rec.vals.view(np.uint64), ref.vals.view(np.uint64))
    assert np.array_equal(red.vals.view(np.uint64), ref.reduced.vals.view(np.uint64))


def gap_lps():
    return [("qap6", G.generate_qap(6, seed=105)),
            ("gen2k", G.generate(2000, 3000, seed=107, degree_mode=0, mean_degree=6.0))]


@pytest.mark.parametrize("name,lp", gap_lps(), ids=[x[0] for x in gap_lps()])
def test_learned_postsolve_objective_gap(name, lp):
    """Reduced LP of the learned schedule, solved by HiGHS and postsolved on the
    GPU, matches HiGHS on the original within 1e-6 (relative), as the plain
    schedule does."""
    blob = weights()
    st0, obj0, *_ = solve(lp)
    assert st0 == 0
    for params in (G.parity_params(), L.learned_params(blob)):
        with G.DeviceLP(lp) as d:
            s = G.preprocess(d, params)
            try:
                st, obj_r, xr, yr, zr = solve(s.reduced())
                assert st == 0
                x, _, _ = G.postprocess(s, xr, yr, zr, d)
            finally:
                s.free()
        assert relative_gap(float(lp.c @ x) + lp.obj_constant, obj0) <= 1e-6
        assert primal_residual(lp, x) <= 1e-6


def test_learned_schedule_gains_over_plain():
    """Held-out seeds of both training families (seeds >= 10^6 are never used
    for training): the learned schedule launches fewer rules and takes less
    presolve time in total than the plain schedule."""
    blob = weights()
    cases = [G.generate_qap(n, seed=1_000_000 + n) for n in (14, 20)]
    cases += [G.generate(200_000, 300_000, seed=1_000_001 + k, degree_mode=0, mean_degree=6.0) for k in range(2)]
    plain_p = G.parity_params(collect_trace=1)
    learned_p = L.learned_params(blob, collect_trace=1)
    totals = {"plain": [0.0, 0], "learned": [0.0, 0]}
    for lp in cases:
        with G.DeviceLP(lp) as d:
            for name, p in (("plain", plain_p), ("learned", learned_p)):
                G.preprocess(d, p).free()  # warm the pool
                best = None
                for _ in range(3):
                    G.check(G.load().gp_synchronize())
                    t = time.perf_counter()
                    s = G.preprocess(d, p)
                    dt = time.perf_counter() - t
                    launches = L.rule_launches(s.trace())
                    s.free()
                    best = dt if best is None else min(best, dt)
                totals[name][0] += best
                totals[name][1] += launches
    print(f"plain {totals['plain'][0] * 1e3:.1f} ms / {totals['plain'][1]} rule launches; "
          f"learned {totals['learned'][0] * 1e3:.1f} ms / {totals['learned'][1]} rule launches")
    assert totals["learned"][1] < totals["plain"][1]
    assert totals["learned"][0] < totals["plain"][0]
