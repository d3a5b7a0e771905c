"""The benchmark configurations at their own sizes on the GPU.

* Config 3 (5M x 8M, ~61M nnz, both the integer recipe of the bench and the
  FP64 recipe): the full presolve -- schedule trace, status, counts, the
  objective constant's bits and every map, tape and reduced-LP array --
  equals the CPU restatement's, compared through the digests
  tests/golden/make_config_digests.py committed.
* Config 4 (64 instances, rows log-uniform in [1e3, 1e5], FP64 data): the
  batched presolve equals the restatement's presolve of the host direct sum,
  through committed digests.
* Config 5 (QAP n=12 and the generator at 200k rows): the learned schedule
  equals the restatement's gated schedule on both.  On QAP n=12 it also
  reduces to exactly the LP the plain schedule reduces to, so both reach the
  same postsolved objective at any accuracy; on the generator LP the two
  schedules stop at different reductions (the learned one skips rules the
  plain one fires), and the objective gap between them needs an LP solve
  that this test does not run.
"""
import json
import os

import numpy as np
import pytest

import gpupre as G
import oracle as O
from digests import compare, digest_engine

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_digests.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["large", "large_fp64"])
def test_config3_full_size_matches_restatement(name):
    exp = golden()["workloads"][name]
    print(f"[config3] {name}: generating {exp['rows']} x {exp['cols']}", flush=True)
    lp = G.generate_workload(name, seed=0)
    assert (lp.m, lp.n, lp.nnz) == (exp["rows"], exp["cols"], exp["nnz"])
    with G.DeviceLP(lp) as d:
        del lp
        s = G.preprocess(d, G.parity_params(collect_trace=1))
        try:
            print(f"[config3] {name}: {s.reduced_rows} x {s.reduced_cols} nnz={s.reduced_nnz} "
                  f"iterations={s.iterations}", flush=True)
            got = digest_engine(s)
        finally:
            s.free()
    compare(got, exp, f"config 3 {name}")


def test_config4_batch_fp64_matches_restatement():
    exp = golden()["batch"]
    print(f"[config4] {exp['count']} instances", flush=True)
    lps = G.batch_workload(exp["count"], seed=exp["seed"], integral=0)
    assert sum(lp.nnz for lp in lps) == exp["nnz"]
    dl = [G.DeviceLP(lp) for lp in lps]
    b = G.preprocess_batch(dl, G.parity_params(collect_trace=1))
    try:
        w = b.whole()
        print(f"[config4] {w.reduced_rows} x {w.reduced_cols} iterations={w.iterations} "
              f"infeasible={w.has_infeasible}", flush=True)
        got = digest_engine(w)
    finally:
        b.free()
        for d in dl:
            d.free()
    compare(got, exp, "config 4 batch")


def learned_blob():
    import learned as L

    if not os.path.exists(L.WEIGHTS_PATH):
        pytest.skip("no trained scheduler weights in the package")
    return L, L.load_blob()


def config5_lp(name):
    if name == "qap12":
        return G.generate_qap(12, seed=211)
    return G.generate(200_000, 300_000, seed=213, degree_mode=0, mean_degree=6.0)


@pytest.mark.parametrize("name", ["qap12", "gen200k"])
def test_config5_learned_schedule_parity(name):
    L, blob = learned_blob()
    lp = config5_lp(name)
    print(f"[config5] {name}: {lp.m} x {lp.n} nnz={lp.nnz}", flush=True)
    lparams = L.learned_params(blob, collect_trace=1)
    ref = O.presolve(lp, lparams)
    out = {}
    with G.DeviceLP(lp) as d:
        for tag, p in (("learned", lparams), ("plain", G.parity_params(collect_trace=1))):
            s = G.preprocess(d, p)
            try:
                out[tag] = (s.record(), s.reduced(), s.trace(), s.obj_constant_delta, s.iterations)
                if tag == "learned":
                    # postsolve of one reduced point through both records below
                    rng = np.random.default_rng(5)
                    red = out[tag][1]
                    x = np.clip(rng.standard_normal(red.n), red.l, red.u)
                    out["x_learned"] = G.postprocess(s, x, np.zeros(red.m), np.zeros(red.n), None)[0]
                    out["x_red"] = x
                else:
                    out["x_plain"] = G.postprocess(s, out["x_red"], np.zeros(s.reduced_rows),
                                                   np.zeros(s.reduced_cols), None)[0]
            finally:
                s.free()
    rec, red, trace, obj, _ = out["learned"]
    # the learned schedule on the GPU is the restatement's gated schedule, bit for bit
    assert [(t.phase, t.fired_mask, t.skipped_mask, t.final_sweep) for t in trace] == \
           [(t.phase, t.fired_mask, t.skipped_mask, t.final_sweep) for t in ref.trace]
    for k in ("row_red2org", "col_red2org", "types", "index_starts", "indices"):
        assert np.array_equal(getattr(rec, k), getattr(ref, k)), k
    assert np.array_equal(rec.vals.view(np.uint64), ref.vals.view(np.uint64))
    for k in ("row_ptr", "col_idx", "c", "l", "u", "AL", "AU"):
        assert np.array_equal(getattr(red, k), getattr(ref.reduced, k)), k
    prec, pred, _, pobj, _ = out["plain"]
    print(f"[config5] {name}: learned -> {red.m} x {red.n}, plain -> {pred.m} x {pred.n}", flush=True)
    if name != "qap12":
        return
    # the plain schedule reduces to the very same LP with the same objective
    # constant: the two postsolved objectives are then equal at any gap
    for k in ("row_ptr", "col_idx", "c", "l", "u", "AL", "AU"):
        assert np.array_equal(getattr(red, k), getattr(pred, k)), f"learned vs plain reduced {k}"
    assert np.array_equal(red.vals.view(np.uint64), pred.vals.view(np.uint64))
    assert np.array_equal(rec.row_red2org, prec.row_red2org) and np.array_equal(rec.col_red2org, prec.col_red2org)
    assert obj == pobj
    c = lp.c
    assert float(c @ out["x_learned"]) == pytest.approx(float(c @ out["x_plain"]), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("n,published", [(60, (432120, 6485400, 26359200)), (70, (686140, 12012350, 48715800))])
def test_qap_reduction_equals_published(n, published):
    """QAP LPs of the shape of the reference's tai60b (n = 60) and lipa70a
    (n = 70) reduce, under the reference's default parameters, to exactly the
    dimensions of its published logs
    (results/raw/benchmarks/QAP20/03_GPU-Preprocessor/logs/
    tables_3_4_5_qap20_gpu_preprocessor_{tai60b,lipa70a}.log:22,27)."""
    lp = G.generate_qap(n, seed=n)
    print(f"[qap published] n={n}: {lp.m} x {lp.n} nnz={lp.nnz}", flush=True)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, G.default_params())
        try:
            assert not s.has_infeasible and not s.has_unbounded
            assert (s.reduced_rows, s.reduced_cols, s.reduced_nnz) == published
            assert s.folding_applied == 0
        finally:
            s.free()
