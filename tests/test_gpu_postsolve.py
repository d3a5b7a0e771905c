"""Postsolve on the device vs the CPU restatement's serial replay.

The engine replays long tapes component by component (records that touch
disjoint columns / rows run on different lanes, each component in reverse
tape order); short tapes replay serially.  Both must give the bits of the
serial replay (gpu_postsolve.cu:149-173 serial and component kernels).
GP_POSTSOLVE_PARALLEL_MIN forces either mode."""
import os

import numpy as np
import pytest

import gpupre as G
import oracle as O
from lp_check import primal_residual, relative_gap, solve
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

MODES = {"serial": str(2**62), "parallel": "0"}


def gpu_postsolve(lp, params, x, y, z, mode, with_model=True):
    old = os.environ.get("GP_POSTSOLVE_PARALLEL_MIN")
    os.environ["GP_POSTSOLVE_PARALLEL_MIN"] = MODES[mode]
    try:
        with G.DeviceLP(lp) as d:
            s = G.preprocess(d, params)
            try:
                return G.postprocess(s, x, y, z, d if with_model else None)
            finally:
                s.free()
    finally:
        if old is None:
            os.environ.pop("GP_POSTSOLVE_PARALLEL_MIN", None)
        else:
            os.environ["GP_POSTSOLVE_PARALLEL_MIN"] = old


def random_solution(res, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(res.n1), rng.standard_normal(res.m1), rng.standard_normal(res.n1)


def check_same(got, exp, label):
    for g, e, k in zip(got, exp, "xyz"):
        assert_same(g, e, f"{label} {k}")


@pytest.mark.parametrize("mode", ["serial", "parallel"])
@pytest.mark.parametrize("cpu_tape", [0, 1])
@pytest.mark.parametrize("seed", [0, 3])
def test_postsolve_small(seed, cpu_tape, mode):
    lp = G.generate_workload("small", seed=seed)
    params = G.parity_params(record_postsolve_tape_cpu=cpu_tape)
    res = O.presolve(lp, params)
    x, y, z = random_solution(res, seed + 17)
    print(f"[postsolve] small seed={seed} cpu_tape={cpu_tape} mode={mode} records={res.types.size}", flush=True)
    check_same(gpu_postsolve(lp, params, x, y, z, mode), O.postsolve(res, lp, x, y, z), "small")


@pytest.mark.parametrize("mode", ["serial", "parallel"])
def test_postsolve_small_without_model(mode):
    lp = G.generate_workload("small", seed=4)
    params = G.parity_params(record_postsolve_tape_cpu=1)
    res = O.presolve(lp, params)
    x, y, z = random_solution(res, 4)
    check_same(gpu_postsolve(lp, params, x, y, z, mode, with_model=False), O.postsolve(res, None, x, y, z),
               "no model")


@pytest.mark.parametrize("cpu_tape", [0, 1])
def test_postsolve_medium_parallel(cpu_tape):
    lp = G.generate_workload("medium", seed=0)
    params = G.parity_params(record_postsolve_tape_cpu=cpu_tape)
    res = O.presolve(lp, params)
    x, y, z = random_solution(res, 11)
    print(f"[postsolve] medium cpu_tape={cpu_tape} records={res.types.size}", flush=True)
    check_same(gpu_postsolve(lp, params, x, y, z, "parallel"), O.postsolve(res, lp, x, y, z), "medium")


def implied_bound_solution(res, lp, seed):
    """A reduced solution that puts kept columns of RowImpliedBoundChange
    records on their implied bound, so those records do update y and z."""
    x, y, z = random_solution(res, seed)
    red_of = np.full(res.n0, -1, np.int64)
    red_of[res.col_red2org] = np.arange(res.n1)
    hits = 0
    for k in np.nonzero(res.types == G.RED_ROW_IMPLIED_BOUND_CHANGE)[0][::2]:
        col = res.indices[res.index_starts[k]]
        ol, ou, nl, nu = res.vals[res.value_starts[k]:res.value_starts[k] + 4]
        lch, uch = nl > ol + 1e-7, nu < ou - 1e-7
        if lch == uch or red_of[col] < 0:
            continue
        x[red_of[col]] = nl if lch else nu
        hits += 1
    return x, y, z, hits


@pytest.mark.parametrize("mode", ["serial", "parallel"])
@pytest.mark.parametrize("workload", ["small", "medium"])
def test_postsolve_implied_bound_records_active(workload, mode):
    lp = G.generate_workload(workload, seed=2)
    params = G.parity_params(record_postsolve_tape_cpu=1)
    res = O.presolve(lp, params)
    x, y, z, hits = implied_bound_solution(res, lp, 5)
    print(f"[postsolve] {workload} implied-bound records placed={hits} mode={mode}", flush=True)
    assert hits > 0
    exp = O.postsolve(res, lp, x, y, z)
    # the placed records do change the dual: same replay without them differs
    x0, y0, z0 = random_solution(res, 5)
    base = O.postsolve(res, lp, x0, y0, z0)
    assert not np.array_equal(exp[1], base[1]) or not np.array_equal(exp[2], base[2])
    check_same(gpu_postsolve(lp, params, x, y, z, mode), exp, f"{workload} implied bounds")


def test_postsolve_large10_parallel_matches_serial():
    """The bench recipe at 1/10 scale: component replay == serial replay."""
    lp = G.generate_workload("large10", seed=0)
    params = G.default_params()
    with G.DeviceLP(lp) as d:
        s = G.preprocess(d, params)
        n1, m1, count = s.reduced_cols, s.reduced_rows, s.record().types.size
        s.free()
    rng = np.random.default_rng(3)
    x, y, z = rng.standard_normal(n1), rng.standard_normal(m1), rng.standard_normal(n1)
    print(f"[postsolve] large10 records={count}", flush=True)
    par = gpu_postsolve(lp, params, x, y, z, "parallel")
    ser = gpu_postsolve(lp, params, x, y, z, "serial")
    check_same(par, ser, "large10")


def test_postsolve_highs_gap_parallel():
    """HiGHS on the reduced Zipf LP, device postsolve in component mode: the
    recovered point is the serial replay's and meets the 1e-6 gap."""
    w = dict(G.WORKLOADS["medium"])
    w.pop("m"), w.pop("n")
    lp = G.generate(4000, 6000, 5, **w)
    params = G.parity_params(record_postsolve_tape_cpu=1)
    st0, obj0, *_ = solve(lp)
    res = O.presolve(lp, params)
    st, obj_r, xr, yr, zr = solve(res.reduced)
    assert st0 == 0 and st == 0
    got = gpu_postsolve(lp, params, xr, yr, zr, "parallel")
    check_same(got, O.postsolve(res, lp, xr, yr, zr), "zipf-4k highs")
    x = got[0]
    assert relative_gap(float(np.dot(lp.c, x)) + lp.obj_constant, obj0) <= 1e-6
    assert primal_residual(lp, x) <= 1e-6
