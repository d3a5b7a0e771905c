"""The reference's postsolve known answers (cpp/tests/test_gpu_postsolve.cpp)
on the CPU restatement: per-record replay formulas, the original-model dual
handlers, structural recovery checkpoints, tape/timeline rejection, and the
records that small presolve runs write."""
import numpy as np
import pytest

import oracle as O
from postsolve_cases import (PRESOLVE_CASES, RECORD_CASES, PartialModel, case_host_lp, check_expect, host_record,
                             params_for, resolve)


@pytest.mark.parametrize("case", RECORD_CASES, ids=[c["name"] for c in RECORD_CASES])
def test_oracle_record_postsolve(case):
    rec = host_record(case)
    model = PartialModel(case.get("original")) if case.get("original") else None
    out = O.postsolve_record(rec, model.host() if model else None, case["x"], case["y"], case["z"])
    if case.get("expect_error"):
        assert out is None, "malformed record accepted"
        return
    assert out is not None, "record rejected"
    for name, got in zip("xyz", out):
        if case["expect"].get(name) is not None:
            check_expect(got, case["expect"][name], case["tol"], f"{case['name']} {name}")


@pytest.mark.parametrize("case", PRESOLVE_CASES, ids=[c["name"] for c in PRESOLVE_CASES])
def test_oracle_presolve_record_postsolve(case):
    lp = case_host_lp(case)
    res = O.presolve(lp, params_for(case))
    er = case["expect_record"]
    assert (res.m1, res.n1) == (er["reduced_rows"], er["reduced_cols"])
    for k in ("m0", "n0"):
        if k in er:
            assert getattr(res, k) == er[k]
    for k in ("col_red2org", "fixed_idx"):
        if k in er:
            assert list(getattr(res, k)) == er[k], k
    if "fixed_val" in er:
        check_expect(res.fixed_val, resolve(er["fixed_val"], case) if isinstance(er["fixed_val"], list)
                     else [resolve(er["fixed_val"], case)], case["tol"], "fixed_val")
    if "types" in er:
        assert list(res.types) == er["types"] and list(res.indices) == er["indices"]
        assert abs(res.vals[0] - er["vals0"]) < 1e-12
    n1, m1 = res.n1, res.m1
    x = case["x"] if case["x"] is not None else np.zeros(n1)
    y = case["y"] if case["y"] is not None else np.zeros(m1)
    z = case["z"] if case["z"] is not None else np.zeros(n1)
    xo, yo, zo = O.postsolve(res, lp if case["original"] else None, x, y, z)
    for name, got in zip("xyz", (xo, yo, zo)):
        if case["expect"].get(name) is not None:
            check_expect(got, resolve(case["expect"][name], case), case["tol"], f"{case['name']} {name}")


from postsolve_cases import paired_fixtures  # noqa: E402

PAIRED = paired_fixtures()


@pytest.mark.parametrize("name,fixture,model", PAIRED, ids=[p[0] for p in PAIRED])
def test_oracle_paired_fixture_replays(name, fixture, model):
    """Every fixture of the reference's serial == parallel test replays on
    the restatement (the engine's rounds replay is compared to it on the
    GPU); some families have a closed form checked here."""
    rec = fixture.record()
    pm = PartialModel(model) if model else None
    x, y, z = fixture.inputs()
    out = O.postsolve_record(rec, pm.host() if pm else None, x, y, z)
    assert out is not None
    xo, yo, zo = out
    count = 2048
    if name == "components_4":
        # InfiniteBoundDualFix: the column takes the smallest value >= 0 that
        # satisfies its recorded row x_e - x_k >= 1
        xe, xk = xo[:count], xo[count:2 * count]
        assert np.all(xe - xk >= 1.0 - 1e-12) and np.all(xe >= 0.0)
        assert np.all((xe == 0.0) | (np.abs(xe - xk - 1.0) <= 1e-12))
    if name == "components_5":
        # FmeCol with one lower side x_e >= -0.25 + 0.5 x_k: the bound is tight
        assert np.array_equal(xo[:count], -0.25 + 0.5 * xo[count:2 * count])
        assert np.all(yo[:count] == 0.0) and np.all(zo[:count] == 0.0)
    if name == "components_0":
        # DoubletonEquation 2 x_e - x_k = 3 holds for every pair
        assert np.allclose(2.0 * xo[:count] - xo[count:2 * count], 3.0, rtol=0, atol=1e-12)
    if name == "subcol_0":
        # singleton column: 2 x_c + 1e16 x_s + 0.25 x_{n-1} - 1e16 x_s = 3, in that summation order
        xs, xl = xo[count], xo[count + 1]
        for i in (0, 1, 7, count - 1):
            partial = 1e16 * xs + 0.25 * xl
            partial = partial + (-1e16) * xs
            assert xo[i] == (3.0 - partial) / 2.0
