"""Rule-level backends for parity tests: the CPU restatement (oracle) and the
HIP engine expose the same plan / stats / rule interface here."""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

import gpupre as G
import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def decode(v):
    if isinstance(v, list):
        return [decode(x) for x in v]
    if v == "inf":
        return np.inf
    if v == "-inf":
        return -np.inf
    return v


def load_kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return [dict((k, decode(v) if k not in ("A", "params", "flags") else v) for k, v in c.items())
                for c in json.load(f)["cases"]]


def case_lp(case) -> G.HostLP:
    A = case["A"]
    if isinstance(A, dict) and "row_ptr" in A:
        m, n = A["m"], A["n"]
        lp = G.HostLP(m, n, np.asarray(A["row_ptr"], np.int32), np.asarray(A["col_idx"], np.int32),
                      np.asarray(decode(A["vals"]), np.float64), np.zeros(n), np.zeros(n), np.zeros(n),
                      np.zeros(m), np.zeros(m))
    elif isinstance(A, dict):
        m, n = A["m"], A["n"]
        lp = G.HostLP(m, n, np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(n),
                      np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m))
    else:
        dense = np.asarray(A, dtype=np.float64)
        m, n = dense.shape
        lp = G.HostLP.from_dense(dense, np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m))
    plan = case.get("plan", {})
    lpv = case.get("lp", {})
    # The reference rule tests leave LPInfoGpu vectors null; the rules then
    # fall back to the plan vectors, so the LP carries the plan's entry state.
    for name, src in (("c", "new_c"), ("l", "new_l"), ("u", "new_u"), ("AL", "new_AL"), ("AU", "new_AU")):
        if name in lpv:
            setattr(lp, name, np.asarray(lpv[name], np.float64))
        elif src in plan:
            setattr(lp, name, np.asarray(plan[src], np.float64))
    return lp


class OracleBackend:
    name = "oracle"

    def __init__(self, lp: G.HostLP, tape=True):
        self.p = O.Plan(lp, tape=tape)

    def set(self, field, values):
        self.p.set(field, values)

    def get(self, field):
        return self.p.get(field)

    def refresh(self):
        self.p.refresh()

    def apply(self, rule, params):
        self.p.apply(rule, params)

    def flags(self):
        return self.p.flags()

    def startup(self, rule, params, pending=True):
        return self.p.startup(rule, params, pending)

    def obj(self):
        return float(self.p.plan.obj_constant_delta)

    def set_flag(self, name, v):
        setattr(self.p.plan, name, v)

    def tape(self):
        return self.p.tape()

    def set_live(self, v):
        self.p.stats.live_masks = v

    def new_A(self):
        """Replacement matrix of the plan (row_ptr, col_idx, vals) or None."""
        A = self.p.plan.new_A
        if not self.p.plan.has_new_A:
            return None
        from oracle import _arr
        return (_arr(A.row_ptr, A.rows + 1, G.NP_OFFSET), _arr(A.col_idx, A.nnz, np.int32),
                _arr(A.vals, A.nnz, np.float64))


class GpuBackend:
    name = "gpu"
    _FIELDS = O.Plan._FIELDS
    _STATS = O.Plan._STATS

    def __init__(self, lp: G.HostLP, tape=True):
        lib = G.load()
        self.m, self.n = lp.m, lp.n
        self.dlp = G.DeviceLP(lp)
        self.plan = G.GpPlan()
        self.stats = G.GpStats()
        G.check(lib.gp_plan_create(C.byref(self.dlp.lp), C.byref(self.plan)))
        G.check(lib.gp_stats_create(lp.m, lp.n, C.byref(self.stats)))
        # stats start zeroed like a default PresolveStatsGpu
        for name, (dim, dt) in self._STATS.items():
            G.to_device(getattr(self.stats, name), np.zeros(getattr(self, dim), dt))
        if tape:
            G.check(lib.gp_plan_enable_tape(C.byref(self.plan)))

    def _ptr(self, field):
        return getattr(self.plan if field in self._FIELDS else self.stats, field)

    def set(self, field, values):
        dim, dt = (self._FIELDS if field in self._FIELDS else self._STATS)[field]
        a = np.asarray(values, dtype=dt)
        assert a.size == getattr(self, dim)
        G.to_device(self._ptr(field), a)

    def get(self, field):
        dim, dt = (self._FIELDS if field in self._FIELDS else self._STATS)[field]
        return G.to_host(self._ptr(field), getattr(self, dim), dt)

    def refresh(self):
        G.check(G.load().gp_refresh_live_stats(C.byref(self.stats), C.byref(self.dlp.lp), C.byref(self.plan)))

    def apply(self, rule, params):
        rid = G.RULE[rule] if isinstance(rule, str) else int(rule)
        G.check(G.load().gp_rule_apply(rid, C.byref(self.plan), C.byref(self.dlp.lp), C.byref(self.stats),
                                       C.byref(params)))

    def flags(self):
        return {k: int(getattr(self.plan, k)) for k in G.PLAN_FLAGS}

    def startup(self, rule, params, pending=True):
        """gp_startup_rule into this plan; returns (applicable, count)."""
        rid = G.STARTUP_RULE[rule] if isinstance(rule, str) else int(rule)
        count, ok = C.c_int32(), C.c_int32()
        G.check(G.load().gp_startup_rule(rid, C.byref(self.dlp.lp), C.byref(params), C.byref(self.plan),
                                         1 if pending else 0, C.byref(count), C.byref(ok)))
        return bool(ok.value), count.value

    def obj(self):
        return float(self.plan.obj_constant_delta)

    def set_flag(self, name, v):
        setattr(self.plan, name, v)

    def set_live(self, v):
        self.stats.live_masks = v

    def new_A(self):
        A = self.plan.new_A
        if not self.plan.has_new_A:
            return None
        return (G.to_host(A.row_ptr, A.rows + 1, G.NP_OFFSET), G.to_host(A.col_idx, A.nnz, np.int32),
                G.to_host(A.vals, A.nnz, np.float64))

    def tape(self):
        if not self.plan.tape:
            return None
        r, i, v = C.c_int64(), C.c_int64(), C.c_int64()
        G.check(G.load().gp_plan_tape_sizes(C.byref(self.plan), C.byref(r), C.byref(i), C.byref(v)))
        out = dict(types=np.empty(r.value, np.int32), index_starts=np.empty(r.value + 1, G.NP_OFFSET),
                   value_starts=np.empty(r.value + 1, G.NP_OFFSET), dual_modes=np.empty(r.value, np.uint8),
                   indices=np.empty(i.value, np.int32), vals=np.empty(v.value, np.float64))
        G.check(G.load().gp_plan_tape_copy(C.byref(self.plan), *[a.ctypes.data_as(C.c_void_p) for a in
                                                                   (out["types"], out["index_starts"],
                                                                    out["value_starts"], out["dual_modes"],
                                                                    out["indices"], out["vals"])]))
        return out

    def free(self):
        G.load().gp_plan_destroy(C.byref(self.plan))
        G.load().gp_stats_destroy(C.byref(self.stats))
        self.dlp.free()


def make_backend(kind, lp, tape=True):
    return OracleBackend(lp, tape) if kind == "oracle" else GpuBackend(lp, tape)


def run_kat(case, kind):
    """Run one transcribed reference case on a backend; returns (backend, state after)."""
    lp = case_lp(case)
    b = make_backend(kind, lp)
    for k, v in case.get("plan", {}).items():
        b.set(k, v)
    for k, v in case.get("stats", {}).items():
        b.set(k, v)
    b.set_live(0)
    params = G.default_params(**case.get("params", {}))
    calls = case.get("calls", 1)
    states = []
    for c in range(calls):
        if c > 0 and case.get("reset_has_change_between_calls"):
            b.set_flag("has_change", 0)
        b.apply(case["rule"], params)
        states.append({k: b.get(k) for k in O.Plan._FIELDS})
    return b, states


def check_kat(case, b, states):
    tol = case.get("tol", 0.0)
    if "after_first" in case:
        for k, v in case["after_first"].items():
            assert_close(states[0][k], v, tol, f"{case['name']} after first call: {k}")
    exp = case["expect"]
    final = states[-1]
    for k, v in exp.items():
        if k == "flags":
            fl = b.flags()
            for fk, fv in v.items():
                assert fl[fk] == fv, f"{case['name']}: flag {fk} = {fl[fk]}, expected {fv}"
        elif k == "obj_constant_delta":
            assert abs(b.obj() - v) <= max(tol, 1e-15), f"{case['name']}: obj {b.obj()} vs {v}"
        elif k == "tape":
            t = b.tape()
            for tk, tv in v.items():
                got = t[tk]
                if tk == "vals":
                    assert_close(got, tv, tol, f"{case['name']}: tape vals")
                else:
                    assert list(got) == list(tv), f"{case['name']}: tape {tk} = {list(got)}, expected {tv}"
        else:
            assert_close(final[k], v, tol, f"{case['name']}: {k}")


def assert_close(got, exp, tol, what):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    assert got.shape == exp.shape, f"{what}: shape {got.shape} vs {exp.shape}"
    with np.errstate(invalid="ignore"):
        same_inf = (np.isinf(got) & np.isinf(exp) & (np.sign(got) == np.sign(exp))) | (np.isnan(got) & np.isnan(exp))
        fin = ~(np.isinf(exp) | np.isnan(exp))
        ok = same_inf | (fin & (np.abs(got - exp) <= tol))
    assert ok.all(), f"{what}: got {got.tolist()}, expected {exp.tolist()}"


def gpu_startup_analyze(lp: G.HostLP, rule, params):
    """gp_startup_rule without a plan (analysis only): (applicable, count)."""
    rid = G.STARTUP_RULE[rule] if isinstance(rule, str) else int(rule)
    count, ok = C.c_int32(), C.c_int32()
    with G.DeviceLP(lp) as d:
        G.check(G.load().gp_startup_rule(rid, C.byref(d.lp), C.byref(params), None, 0, C.byref(count),
                                         C.byref(ok)))
    return bool(ok.value), count.value
