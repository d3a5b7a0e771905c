"""Postsolve known answers transcribed from the reference's tests
(tests/golden/reference_postsolve.json) and the builders that turn a case into
a gp_host_record, a partial original model and a host LP."""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

import gpupre as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_postsolve.json")


def decode(v):
    if isinstance(v, list):
        return [decode(x) for x in v]
    if isinstance(v, dict):
        return {k: decode(x) for k, x in v.items()}
    if v == "inf":
        return np.inf
    if v == "-inf":
        return -np.inf
    return v


def load():
    with open(GOLDEN) as f:
        d = json.load(f)
    return [decode(c) for c in d["cases"]], [decode(c) for c in d["presolve_cases"]]


RECORD_CASES, PRESOLVE_CASES = load()


def host_record(case) -> G.HostRecord:
    r = case["record"]
    t = r.get("tape", {})
    return G.HostRecord(r["m0"], r["n0"], r["m1"], r["n1"], r.get("row_red2org", []), r.get("col_red2org", []),
                        r.get("fixed_idx", []), r.get("fixed_val", []), t.get("types", []),
                        t.get("index_starts", [0]), t.get("value_starts", [0]), t.get("dual_modes", []),
                        t.get("indices", []), t.get("vals", []), r.get("abs_value", []), r.get("var_pair", []),
                        [tuple(c) for c in r.get("checkpoints", [])])


class PartialModel:
    """The LPInfoGpu a reference test passes: only the fields it sets, the
    rest null.  host() gives a gp_device_lp over host arrays (oracle), device()
    one over device copies (engine)."""

    VEC = ("c", "AL", "AU", "l", "u")

    def __init__(self, spec):
        self.spec = spec or {}
        self._host = None
        self._dev = None
        self._keep = []

    def _arrays(self, mat):
        return (np.asarray(mat["row_ptr"], G.NP_OFFSET), np.asarray(mat["col_idx"], np.int32),
                np.asarray(mat["vals"], np.float64))

    def host(self) -> G.GpDeviceLp:
        if self._host is None:
            lp = G.GpDeviceLp()
            for name in ("A", "AT"):
                if name in self.spec:
                    m = self.spec[name]
                    rp, ci, va = self._arrays(m)
                    self._keep += [rp, ci, va]
                    setattr(lp, name, G.GpCsr(m["rows"], m["cols"], len(ci), rp.ctypes.data_as(C.c_void_p),
                                              ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p)))
            for name in self.VEC:
                if name in self.spec:
                    a = np.asarray(self.spec[name], np.float64)
                    self._keep.append(a)
                    setattr(lp, name, a.ctypes.data_as(C.c_void_p))
            self._host = lp
        return self._host

    def device(self) -> G.GpDeviceLp:
        if self._dev is None:
            lp = G.GpDeviceLp()
            for name in ("A", "AT"):
                if name in self.spec:
                    m = self.spec[name]
                    d = [G.DeviceArray.from_host(a) for a in self._arrays(m)]
                    self._keep += d
                    setattr(lp, name, G.GpCsr(m["rows"], m["cols"], len(m["col_idx"]), d[0].ptr, d[1].ptr, d[2].ptr))
            for name in self.VEC:
                if name in self.spec:
                    d = G.DeviceArray.from_host(np.asarray(self.spec[name], np.float64))
                    self._keep.append(d)
                    setattr(lp, name, d.ptr)
            self._dev = lp
        return self._dev


def case_host_lp(case) -> G.HostLP:
    s = case["lp"]
    m, n = s["rows"], s["cols"]
    u = np.asarray(s["u"], np.float64)
    if "width" in case:  # upper = lower + width, formed in binary64 as the reference does
        u = u.copy()
        u[0] = float(s["l"][0]) + case["width"]
    vec = lambda k, size: np.asarray(s[k] if s.get(k) is not None else np.zeros(size), np.float64)
    return G.HostLP(m, n, np.asarray(s["row_ptr"], G.NP_OFFSET), np.asarray(s["col_idx"], np.int32),
                    np.asarray(s["vals"], np.float64), vec("c", n), vec("l", n), u, vec("AL", m), vec("AU", m))


def midpoint(case) -> float:
    lo = float(case["lp"]["l"][0])
    hi = lo + case["width"]
    return lo + 0.5 * (hi - lo)


def resolve(v, case):
    if isinstance(v, list):
        return [resolve(x, case) for x in v]
    if v == "midpoint":
        return midpoint(case)
    return v


def check_expect(got, expect, tol, what):
    """The reference's assertion: |got - expected| < tol (tol > 0) or exact
    equality (tol == 0); entries given as None are not asserted."""
    got = np.asarray(got)
    assert len(got) == len(expect), f"{what}: length {len(got)} vs {len(expect)}"
    for k, e in enumerate(expect):
        if e is None:
            continue
        if tol == 0:
            assert got[k] == e, f"{what}[{k}] = {got[k]!r}, expected {e!r}"
        else:
            assert abs(got[k] - e) < tol, f"{what}[{k}] = {got[k]!r}, expected {e!r} within {tol}"


def params_for(case) -> G.GpParams:
    p = G.parity_params(**case["params"])
    return p


# ---- serial == parallel replay fixtures (cpp/tests/test_parallel_postsolve.cpp:37-195) ----
# The reference compares its serial replay with its optimized (component /
# independent-range) replay bit for bit on these tapes.  Here the CPU
# restatement replays serially and the engine replays in dependency rounds.
EXACT = 1


class Fixture:
    """Identity maps over n columns and m rows with the reference fixture's
    reduced point; add() appends a record in the Exact dual mode."""

    def __init__(self, n, m=0):
        self.n, self.m = n, m
        i = np.arange(n)
        self.x = 1.0 + (i % 11) * 0.125
        self.z = (i % 5 - 2) * 0.25
        self.y = (np.arange(m) % 7 - 3) * 0.5
        self.types, self.modes, self.idx, self.vals = [], [], [], []
        self.istarts, self.vstarts = [0], [0]
        self.abs_value, self.checkpoints = [], []
        self.null_duals = False

    def add(self, typ, ids, values):
        self.types.append(typ)
        self.modes.append(EXACT)
        self.idx += list(ids)
        self.vals += list(values)
        self.istarts.append(len(self.idx))
        self.vstarts.append(len(self.vals))

    def record(self) -> G.HostRecord:
        return G.HostRecord(self.m, self.n, self.m, self.n, np.arange(self.m), np.arange(self.n), [], [],
                            self.types, self.istarts, self.vstarts, self.modes, self.idx, self.vals,
                            self.abs_value, [], self.checkpoints)

    def inputs(self):
        return (self.x, None, None) if self.null_duals else (self.x, self.y, self.z)


INF = np.inf


def paired_fixtures():
    """(name, fixture, model spec or None) for every fixture family."""
    out = []
    for variant in range(8):
        if variant in (4, 5, 6):  # device-resident segments / scratch limit: reference-internal paths
            continue
        count = 1023 if variant == 1 else 2048
        f = Fixture(count + 2, 2)
        for i in range(count):
            col = count - 1 - i if variant == 2 else i
            if variant == 3 and i == count - 1:
                col = i - 1
            fixed = -0.0 if i % 2 else 0.125
            f.add(G.RED_FIXED_VARIABLE, [col, 0, 1, 0], [fixed, 3.0, 1e16, -0.5, -1e16])
        f.null_duals = variant == 7
        out.append((f"fixed_{variant}", f, None))
    for variant in range(10):
        if variant in (6, 7):
            continue
        count = 2048
        f = Fixture(count + 2, count)
        for i in range(count):
            col, row, support = (count - 1 - i if variant == 1 else i), i, count
            if variant == 2 and i == count - 1:
                col = i - 1
            if variant == 3 and i == count - 1:
                row = i - 1
            if variant == 4 and i == 0:
                support = 1
            if variant == 5 and i == count - 1:
                support = 0
            f.add(G.RED_SINGLETON_COLUMN_SUBSTITUTION, [col, row, 3, support, count + 1, support],
                  [2.0, 3.0, -INF, INF, 0.5, float(i % 2), 1e16, 0.25, -1e16])
        if variant == 8:
            f.abs_value.append({"splits": [(count, count + 1, 1.0)]})
            f.checkpoints.append((G.RECOVERY_ABS_VALUE, 0, 1024))
        f.null_duals = variant == 9
        out.append((f"subcol_{variant}", f, None))
    for variant in range(10):
        if variant in (6, 7):
            continue
        count = 2048
        f = Fixture(2 * count + 2, count)
        for i in range(count):
            elim, keep = i, ((i + 1) % count if variant == 1 else count + i)
            if variant == 2 and i % 2:
                f.add(G.RED_ROW_DUAL_SHIFT, [i], [0.125])
            elif variant == 3:
                f.add(G.RED_PARALLEL_COLUMN, [elim, keep], [1.0, 0.0, INF, 0.0, INF])
            elif variant == 4:
                f.add(G.RED_INFINITE_BOUND_DUAL_FIX, [1, elim, 1, 2, elim, keep], [0.0, 1.0, 1.0, -1.0])
            elif variant == 5:
                f.add(G.RED_FME_COL, [1, elim, i, 1, 1, 1, 1, keep], [-0.25, 0.5])
            else:
                f.add(G.RED_DOUBLETON_EQUATION, [elim, keep, i], [2.0, -1.0, 3.0, -INF, INF, 0.0, INF, 0.0, INF, 0.5])
        if variant == 8:
            f.add(G.RED_ROW_IMPLIED_BOUND_CHANGE, [0, 0], [0.0, 1.0, 0.0, 0.5])
        f.null_duals = variant == 9
        out.append((f"components_{variant}", f, None))
    count = 2048
    ap = [0] + [count] * count
    atp = [min(j, count) for j in range(2 * count + 1)]
    model = {"A": {"rows": count, "cols": 2 * count, "row_ptr": ap, "col_idx": list(range(count)),
                   "vals": [0.25] * count},
             "AT": {"rows": 2 * count, "cols": count, "row_ptr": atp, "col_idx": [0] * count, "vals": [0.25] * count},
             "c": [1.0] * (2 * count), "l": [0.0] * (2 * count), "u": [INF] * (2 * count)}
    f = Fixture(2 * count, count)
    for i in range(count):
        f.add(G.RED_DOUBLETON_EQUATION, [i, count + i, i], [1, -1, 0, -INF, INF, 0, INF, 0, INF, 1])
    out.append(("original_at", f, model))
    f = Fixture(2 * count, count)
    for i in range(count):
        f.add(G.RED_DELETED_ROW, [i, i], [0, INF, 1, 0, 0, 0, 0, 0, 0])
    out.append(("deleted_original_at", f, model))
    f = Fixture(2 * count, count)
    for i in range(count):
        f.add(G.RED_PARALLEL_ROW, [i, (i + 1) % count], [1, 0, INF, -INF, 1])
    out.append(("parallel_rows_original_a", f, model))
    return out
