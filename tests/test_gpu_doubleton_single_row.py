"""The single-row doubleton path (doubleton_equations_scan off) on the GPU:
bit-identical to the CPU restatement on the hand-worked cases, on random
rule states and through whole presolves."""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from backends import GpuBackend, OracleBackend
from test_gpu_parity import assert_same, compare_pipeline, random_state

pytestmark = pytest.mark.gpu

INF = np.inf

CASES = {
    "one_per_call": ([[1, 2, 0, 0], [0, 0, 1, 1], [1, 0, 1, 1]], [1, 3, 0, 0], [0] * 4, [5, 10, 9, 9],
                     [4, 3, -INF], [4, 3, 10]),
    "cancel": ([[1, -1, 0], [0, 2, 1], [3, -3, 1]], [0, 0, 1], [0, -INF, 0], [10, INF, 10], [1, -INF, -INF],
               [1, 8, 9]),
    "insert": ([[1, 1, 0], [1, 0, 1], [0, 1, 1], [0, 1, 0]], [1, 0, 0], [0] * 3, [5] * 3, [2, -INF, -INF, -INF],
               [2, 5, 5, 5]),
    "pivot": ([[1, 2000, 0, 0], [0, 0, 1, 1]], [0] * 4, [0] * 4, [1] * 4, [1, 1], [1, 1]),
    "infeasible": ([[1, 1], [1, 0]], [0, 0], [0, 0], [1, 1], [10, -INF], [10, 5]),
}


def params():
    return G.parity_params(doubleton_equations_scan=0, record_postsolve_tape_cpu=1)


def same_state(gb, ob, what):
    for f in O.Plan._FIELDS:
        assert_same(gb.get(f), ob.get(f), f"{what}: {f}")
    assert gb.flags() == ob.flags(), (what, gb.flags(), ob.flags())
    assert_same(np.array([gb.obj()]), np.array([ob.obj()]), f"{what}: obj delta")
    tg, to = gb.tape(), ob.tape()
    for k in to:
        assert_same(tg[k], to[k], f"{what}: tape {k}")
    ag, ao = gb.new_A(), ob.new_A()
    assert (ag is None) == (ao is None), what
    if ao is not None:
        for x, y, k in zip(ag, ao, ("row_ptr", "col_idx", "vals")):
            assert_same(x, y, f"{what}: new_A {k}")


@pytest.mark.parametrize("name", list(CASES))
def test_hand_cases_match_restatement(name):
    A, c, l, u, AL, AU = CASES[name]
    print(f"[doubleton single row] {name}", flush=True)
    lp = G.HostLP.from_dense(np.asarray(A, float), *[np.asarray(v, float) for v in (c, l, u, AL, AU)])
    ob, gb = OracleBackend(lp), GpuBackend(lp)
    try:
        for b in (ob, gb):
            b.refresh()
            b.apply("doubleton_equations", params())
        same_state(gb, ob, name)
    finally:
        gb.free()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_state_matches_restatement(seed):
    print(f"[doubleton single row] random state seed={seed}", flush=True)
    lp = G.generate_workload("small", seed=seed)
    kr, kc = random_state(lp, 300 + seed)
    ob, gb = OracleBackend(lp), GpuBackend(lp)
    try:
        for b in (ob, gb):
            b.set("keep_row_mask", kr)
            b.set("keep_col_mask", kc)
            b.refresh()
        for call in range(3):  # one elimination per call, on the plan's rewritten A
            for b in (ob, gb):
                b.set_flag("has_change", 0)
                b.apply("doubleton_equations", params())
            same_state(gb, ob, f"seed {seed} call {call}")
    finally:
        gb.free()


@pytest.mark.parametrize("name,seed", [("small", 0), ("small", 1), ("medium", 0)])
def test_pipeline_matches_restatement(name, seed):
    lp = G.generate_workload(name, seed=seed)
    s, ores, rec = compare_pipeline(lp, G.parity_params(doubleton_equations_scan=0), f"{name} single-row seed={seed}")
    try:
        assert (rec.types == G.RED_DOUBLETON_EQUATION).any(), "no single-row elimination ran"
    finally:
        s.free()
