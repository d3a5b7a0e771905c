"""The reference's postsolve known answers (cpp/tests/test_gpu_postsolve.cpp)
on the HIP engine through the C ABI: each case must meet the reference's
expected values and equal the CPU restatement bit for bit."""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from postsolve_cases import (PRESOLVE_CASES, RECORD_CASES, PartialModel, case_host_lp, check_expect, host_record,
                             params_for, resolve)
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", RECORD_CASES, ids=[c["name"] for c in RECORD_CASES])
def test_gpu_record_postsolve(case):
    print(f"[postsolve kat] {case['name']}", flush=True)
    rec = host_record(case)
    model = PartialModel(case.get("original")) if case.get("original") else None
    if case.get("expect_error"):
        with pytest.raises(G.GpError) as e:
            G.postsolve_record(rec, case["x"], case["y"], case["z"], model.device() if model else None)
        assert e.value.code == 3  # GP_ERR_RUNTIME, the reference's std::runtime_error
        assert O.postsolve_record(rec, None, case["x"], case["y"], case["z"]) is None
        return
    got = G.postsolve_record(rec, case["x"], case["y"], case["z"], model.device() if model else None)
    ref = O.postsolve_record(rec, model.host() if model else None, case["x"], case["y"], case["z"])
    for name, g, o in zip("xyz", got, ref):
        if case["expect"].get(name) is not None:
            check_expect(g, case["expect"][name], case["tol"], f"{case['name']} {name}")
        assert_same(g, o, f"{case['name']} {name} vs oracle")


@pytest.mark.parametrize("case", PRESOLVE_CASES, ids=[c["name"] for c in PRESOLVE_CASES])
def test_gpu_presolve_record_postsolve(case):
    print(f"[presolve+postsolve kat] {case['name']}", flush=True)
    lp = case_host_lp(case)
    params = params_for(case)
    res = O.presolve(lp, params)
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        try:
            rec = s.record()
            assert (s.reduced_rows, s.reduced_cols) == (res.m1, res.n1)
            er = case["expect_record"]
            assert (s.reduced_rows, s.reduced_cols) == (er["reduced_rows"], er["reduced_cols"])
            for k in ("row_red2org", "col_red2org", "fixed_idx", "fixed_val", "types", "index_starts",
                      "value_starts", "indices", "vals"):
                assert_same(getattr(rec, k), getattr(res, k), k)
            n1, m1 = res.n1, res.m1
            x = case["x"] if case["x"] is not None else np.zeros(n1)
            y = case["y"] if case["y"] is not None else np.zeros(m1)
            z = case["z"] if case["z"] is not None else np.zeros(n1)
            got = G.postprocess(s, x, y, z, d if case["original"] else None)
        finally:
            s.free()
    ref = O.postsolve(res, lp if case["original"] else None, x, y, z)
    for name, g, o in zip("xyz", got, ref):
        if case["expect"].get(name) is not None:
            check_expect(g, resolve(case["expect"][name], case), case["tol"], f"{case['name']} {name}")
        assert_same(g, o, f"{case['name']} {name} vs oracle")


from postsolve_cases import paired_fixtures  # noqa: E402

PAIRED = paired_fixtures()


@pytest.mark.parametrize("name,fixture,model", PAIRED, ids=[p[0] for p in PAIRED])
def test_gpu_paired_fixture_replays(name, fixture, model, monkeypatch):
    """cpp/tests/test_parallel_postsolve.cpp:37-195: the dependency-round
    replay and the one-lane serial replay of the engine give the CPU
    restatement's serial result bit for bit."""
    print(f"[paired replay] {name}", flush=True)
    rec = fixture.record()
    pm = PartialModel(model) if model else None
    x, y, z = fixture.inputs()
    ref = O.postsolve_record(rec, pm.host() if pm else None, x, y, z)
    for mode, threshold in (("rounds", "1"), ("serial", str(1 << 40))):
        monkeypatch.setenv("GP_POSTSOLVE_PARALLEL_MIN", threshold)
        got = G.postsolve_record(rec, x, y, z, pm.device() if pm else None)
        for k, g, o in zip("xyz", got, ref):
            assert_same(g, o, f"{name} {mode} {k}")
