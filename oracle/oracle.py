"""ctypes wrapper of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker of the HIP engine.  Only the
tests, the smoke check and the bench's CPU-baseline leg may import it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "gpu-preprocessor-ijoc_amd"))

import gpupre as G  # noqa: E402  (struct layouts only)

LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrTape(C.Structure):
    _fields_ = [("count", C.c_int64), ("icount", C.c_int64), ("vcount", C.c_int64), ("cap", C.c_int64),
                ("icap", C.c_int64), ("vcap", C.c_int64), ("types", C.c_void_p), ("istarts", C.c_void_p),
                ("vstarts", C.c_void_p), ("modes", C.c_void_p), ("idx", C.c_void_p), ("vals", C.c_void_p)]


class OrResult(C.Structure):
    _fields_ = [("reduced", G.GpDeviceLp), ("m0", C.c_int32), ("n0", C.c_int32), ("m1", C.c_int32),
                ("n1", C.c_int32), ("row_red2org", C.c_void_p), ("col_red2org", C.c_void_p),
                ("fixed_count", C.c_int32), ("fixed_cap", C.c_int32), ("fixed_idx", C.c_void_p),
                ("fixed_val", C.c_void_p), ("tape", OrTape), ("iterations", C.c_int32),
                ("stopped_at_iteration_limit", C.c_int32), ("has_change", C.c_int32),
                ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32),
                ("obj_constant_delta", C.c_double), ("trace_count", C.c_int32), ("trace_cap", C.c_int32),
                ("trace", C.POINTER(G.GpRoundTrace)), ("primal_only", C.c_int32)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "or_presolve": (C.c_int, [C.POINTER(G.GpDeviceLp), C.POINTER(G.GpParams), C.POINTER(OrResult)]),
            "or_result_free": (None, [C.POINTER(OrResult)]),
            "or_postsolve": (C.c_int, [C.POINTER(OrResult), C.POINTER(G.GpDeviceLp), vp, vp, vp, vp, vp, vp]),
            "or_postsolve_record": (C.c_int, [C.POINTER(G.GpHostRecord), C.POINTER(G.GpDeviceLp), vp, vp, vp, vp,
                                              vp, vp, vp]),
            "or_plan_create": (C.c_int, [C.POINTER(G.GpDeviceLp), C.POINTER(G.GpPlan)]),
            "or_plan_destroy": (None, [C.POINTER(G.GpPlan)]),
            "or_plan_enable_tape": (None, [C.POINTER(G.GpPlan)]),
            "or_stats_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(G.GpStats)]),
            "or_stats_destroy": (None, [C.POINTER(G.GpStats)]),
            "or_refresh_live_stats": (None, [C.POINTER(G.GpStats), C.POINTER(G.GpDeviceLp), C.POINTER(G.GpPlan)]),
            "or_rule_apply": (C.c_int, [C.c_int32, C.POINTER(G.GpPlan), C.POINTER(G.GpDeviceLp),
                                        C.POINTER(G.GpStats), C.POINTER(G.GpParams)]),
            "or_row_activity_summary": (None, [C.POINTER(G.GpCsr), vp, vp, C.c_double, vp, vp, vp, vp, vp, vp]),
            "or_det_sum": (C.c_double, [vp, C.c_int64]),
            "or_tape_init": (None, [C.POINTER(OrTape)]),
            "or_parallel_row_hashes": (None, [C.POINTER(G.GpCsr), vp, vp, C.c_double, vp]),
            "or_tape_free": (None, [C.POINTER(OrTape)]),
            "or_fixed_variable_tape": (C.c_double, [C.POINTER(OrTape), vp, vp, vp, vp, vp, C.POINTER(G.GpCsr),
                                                    C.c_int32, C.c_int]),
            "or_row_sum_finite": (C.c_double, [vp, C.c_int64]),
            "or_transpose": (None, [C.POINTER(G.GpCsr), C.POINTER(G.GpCsr)]),
            "or_csr_free": (None, [C.POINTER(G.GpCsr)]),
            "or_rule_group": (C.c_int, [C.c_int32]),
            "or_schedule_features": (None, [C.POINTER(G.GpRoundTrace), C.c_int32, C.c_int32, C.c_int64,
                                            C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]),
            "or_mlp_forward": (None, [vp, vp, C.c_int32, vp]),
            "or_mlp_loss_grad": (None, [vp, vp, vp, vp, C.c_int32, C.POINTER(C.c_double), vp]),
            "or_mlp_adam_step": (None, [vp, vp, vp, vp, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float]),
            "or_fold_lp": (None, [C.POINTER(G.GpDeviceLp), C.c_double, C.POINTER(OrFold)]),
            "or_fold_free": (None, [C.POINTER(OrFold)]),
            "or_startup_rule": (C.c_int, [C.c_int32, C.POINTER(G.GpDeviceLp), C.POINTER(G.GpParams),
                                          C.POINTER(G.GpPlan), C.c_int, C.POINTER(C.c_int32)]),
            "or_unfold": (None, [C.POINTER(OrFold), vp, vp, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _arr(ptr, count, dtype):
    count = int(count)
    if count == 0 or not ptr:
        return np.empty(count, dtype=dtype)
    ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()


class OrFold(C.Structure):
    _fields_ = [("applied", C.c_int32), ("refinement_ok", C.c_int32), ("rounds", C.c_int32), ("m", C.c_int32),
                ("n", C.c_int32), ("mf", C.c_int32), ("nf", C.c_int32), ("row_class", C.c_void_p),
                ("col_class", C.c_void_p), ("row_scale", C.c_void_p), ("col_scale", C.c_void_p),
                ("lp", G.GpDeviceLp)]


class HostLPView:
    """gp_device_lp over host numpy arrays (kept alive by this object)."""

    def __init__(self, lp: "G.HostLP"):
        atp, ati, atv = lp.transpose()
        self._keep = [np.ascontiguousarray(a) for a in
                      (lp.row_ptr, lp.col_idx, lp.vals, atp, ati, atv, lp.c, lp.AL, lp.AU, lp.l, lp.u)]
        rp, ci, va, tp, ti, tv, c, al, au, l, u = self._keep
        self.lp = G.GpDeviceLp()
        self.lp.A = G.GpCsr(lp.m, lp.n, lp.nnz, _p(rp), _p(ci), _p(va))
        self.lp.AT = G.GpCsr(lp.n, lp.m, lp.nnz, _p(tp), _p(ti), _p(tv))
        self.lp.c, self.lp.AL, self.lp.AU, self.lp.l, self.lp.u = _p(c), _p(al), _p(au), _p(l), _p(u)
        self.lp.obj_constant = float(lp.obj_constant)


class Result:
    def __init__(self, r: OrResult):
        self.m0, self.n0, self.m1, self.n1 = r.m0, r.n0, r.m1, r.n1
        self.row_red2org = _arr(r.row_red2org, r.m1, np.int32)
        self.col_red2org = _arr(r.col_red2org, r.n1, np.int32)
        self.fixed_idx = _arr(r.fixed_idx, r.fixed_count, np.int32)
        self.fixed_val = _arr(r.fixed_val, r.fixed_count, np.float64)
        t = r.tape
        self.types = _arr(t.types, t.count, np.int32)
        self.index_starts = _arr(t.istarts, t.count + 1, G.NP_OFFSET)
        self.value_starts = _arr(t.vstarts, t.count + 1, G.NP_OFFSET)
        self.dual_modes = _arr(t.modes, t.count, np.uint8)
        self.indices = _arr(t.idx, t.icount, np.int32)
        self.vals = _arr(t.vals, t.vcount, np.float64)
        self.iterations = r.iterations
        self.stopped_at_iteration_limit = r.stopped_at_iteration_limit
        self.has_change = r.has_change
        self.has_infeasible = r.has_infeasible
        self.has_unbounded = r.has_unbounded
        self.obj_constant_delta = r.obj_constant_delta
        self.primal_only = r.primal_only
        self.trace = [G.GpRoundTrace.from_buffer_copy(r.trace[i]) for i in range(r.trace_count)]
        d = r.reduced
        m, n, nnz = d.A.rows, d.A.cols, d.A.nnz
        self.reduced = G.HostLP(m, n, _arr(d.A.row_ptr, m + 1, G.NP_OFFSET), _arr(d.A.col_idx, nnz, np.int32),
                                _arr(d.A.vals, nnz, np.float64), _arr(d.c, n, np.float64),
                                _arr(d.l, n, np.float64), _arr(d.u, n, np.float64), _arr(d.AL, m, np.float64),
                                _arr(d.AU, m, np.float64), float(d.obj_constant))
        self.reduced.at = (_arr(d.AT.row_ptr, n + 1, G.NP_OFFSET), _arr(d.AT.col_idx, nnz, np.int32),
                           _arr(d.AT.vals, nnz, np.float64))
        self._raw = r

    def free(self):
        if self._raw is not None:
            load().or_result_free(C.byref(self._raw))
            self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FoldResult:
    """or_fold_lp: the restatement of the reference's folding."""

    def __init__(self, lp: "G.HostLP", tolerance: float = 1e-8):
        view = HostLPView(lp)
        self._f = OrFold()
        load().or_fold_lp(C.byref(view.lp), float(tolerance), C.byref(self._f))
        f = self._f
        self.applied, self.refinement_ok, self.rounds = bool(f.applied), bool(f.refinement_ok), f.rounds
        self.m, self.n, self.mf, self.nf = f.m, f.n, f.mf, f.nf
        have = f.m > 0 and f.n > 0
        self.row_class = _arr(f.row_class, f.m if have else 0, np.int32)
        self.col_class = _arr(f.col_class, f.n if have else 0, np.int32)
        self.row_scale = _arr(f.row_scale, f.m if self.applied else 0, np.float64)
        self.col_scale = _arr(f.col_scale, f.n if self.applied else 0, np.float64)
        self.folded = None
        if self.applied:
            d = f.lp
            mf, nf, nnz = d.A.rows, d.A.cols, d.A.nnz
            self.folded = G.HostLP(mf, nf, _arr(d.A.row_ptr, mf + 1, G.NP_OFFSET), _arr(d.A.col_idx, nnz, np.int32),
                                   _arr(d.A.vals, nnz, np.float64), _arr(d.c, nf, np.float64),
                                   _arr(d.l, nf, np.float64), _arr(d.u, nf, np.float64), _arr(d.AL, mf, np.float64),
                                   _arr(d.AU, mf, np.float64), float(d.obj_constant))
            self.folded.at = (_arr(d.AT.row_ptr, nf + 1, G.NP_OFFSET), _arr(d.AT.col_idx, nnz, np.int32),
                              _arr(d.AT.vals, nnz, np.float64))

    def unfold(self, x_red, y_red, z_red):
        x, y, z = np.empty(self.n), np.empty(self.m), np.empty(self.n)
        xr, yr, zr = (np.ascontiguousarray(v, dtype=np.float64) for v in (x_red, y_red, z_red))
        load().or_unfold(C.byref(self._f), _p(xr), _p(yr), _p(zr), _p(x), _p(y), _p(z))
        return x, y, z

    def __del__(self):
        try:
            load().or_fold_free(C.byref(self._f))
        except Exception:
            pass


def fold(lp: "G.HostLP", tolerance: float = 1e-8) -> FoldResult:
    return FoldResult(lp, tolerance)


def startup_analyze(lp: "G.HostLP", rule, params):
    """or_startup_rule without a plan (analysis only): (applicable, count)."""
    view = HostLPView(lp)
    rid = G.STARTUP_RULE[rule] if isinstance(rule, str) else int(rule)
    count = C.c_int32()
    rc = load().or_startup_rule(rid, C.byref(view.lp), C.byref(params), None, 0, C.byref(count))
    assert rc >= 0
    return bool(rc), count.value


def presolve(lp: "G.HostLP", params) -> Result:
    view = HostLPView(lp)
    r = OrResult()
    load().or_presolve(C.byref(view.lp), C.byref(params), C.byref(r))
    return Result(r)


def postsolve(res: Result, original: "G.HostLP | None", x, y, z):
    view = HostLPView(original) if original is not None else None
    xo, yo, zo = np.empty(res.n0), np.empty(res.m0), np.empty(res.n0)
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
    load().or_postsolve(C.byref(res._raw), C.byref(view.lp) if view else None, _p(x), _p(y), _p(z),
                        _p(xo), _p(yo), _p(zo))
    return xo, yo, zo


def postsolve_record(rec: "G.HostRecord", original: "G.GpDeviceLp | None", x, y, z, original_l=None):
    """or_postsolve_record over a host record; `original` is a gp_device_lp of
    HOST arrays (fields may be null).  Returns (x, y, z) or None when the
    record is rejected as malformed."""
    r = rec.rec
    xo, yo, zo = np.empty(r.n0), np.empty(r.m0), np.empty(r.n0)
    ins = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None for v in (x, y, z)]
    ol = np.ascontiguousarray(original_l, dtype=np.float64) if original_l is not None else None
    rc = load().or_postsolve_record(C.byref(r), C.byref(original) if original is not None else None,
                                    _p(ol) if ol is not None else None,
                                    *[_p(v) if v is not None else None for v in ins], _p(xo), _p(yo), _p(zo))
    if rc != 0:
        return None
    return xo, yo, zo


def fixed_variable_tape(AT, fixed, excluded, values, keep_row, costs, with_tape=True, objective=True):
    """or_fixed_variable_tape on host arrays; AT = (rows, cols, row_ptr,
    col_idx, vals).  Returns (objective, tape dict or None)."""
    rows, cols, rp, ci, va = AT
    keep = [np.ascontiguousarray(a, dt) for a, dt in ((rp, G.NP_OFFSET), (ci, np.int32), (va, np.float64))]
    at = G.GpCsr(rows, cols, len(keep[1]), _p(keep[0]), _p(keep[1]), _p(keep[2]))
    arrs = [None if a is None else np.ascontiguousarray(a, dt)
            for a, dt in ((fixed, np.uint8), (excluded, np.uint8), (values, np.float64), (keep_row, np.uint8),
                          (costs, np.float64))]
    t = OrTape()
    load().or_tape_init(C.byref(t))
    try:
        obj = load().or_fixed_variable_tape(C.byref(t) if with_tape else None,
                                            *[_p(a) if a is not None else None for a in arrs], C.byref(at), rows,
                                            1 if objective else 0)
        tape = None
        if with_tape:
            tape = dict(types=_arr(t.types, t.count, np.int32), index_starts=_arr(t.istarts, t.count + 1, G.NP_OFFSET),
                        value_starts=_arr(t.vstarts, t.count + 1, G.NP_OFFSET),
                        dual_modes=_arr(t.modes, t.count, np.uint8), indices=_arr(t.idx, t.icount, np.int32),
                        vals=_arr(t.vals, t.vcount, np.float64))
        return obj, tape
    finally:
        load().or_tape_free(C.byref(t))


def parallel_row_hashes(m, n, row_ptr, col_idx, vals, keep_row, keep_col=None, tol=1e-12):
    keep = [np.ascontiguousarray(a, dt) for a, dt in ((row_ptr, G.NP_OFFSET), (col_idx, np.int32),
                                                     (vals, np.float64))]
    A = G.GpCsr(m, n, len(keep[1]), _p(keep[0]), _p(keep[1]), _p(keep[2]))
    kr = np.ascontiguousarray(keep_row, np.uint8)
    kc = np.ascontiguousarray(keep_col, np.uint8) if keep_col is not None else None
    out = np.empty(m, np.uint64)
    load().or_parallel_row_hashes(C.byref(A), _p(kr), _p(kc) if kc is not None else None, float(tol), _p(out))
    return out


def det_sum(x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return load().or_det_sum(_p(x), x.size)


def row_sum(x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return load().or_row_sum_finite(_p(x), x.size)


def activity_summary(lp: "G.HostLP", l, u, zero_tol, keep_row=None, keep_col=None):
    view = HostLPView(lp)
    m = lp.m
    mn, mx = np.empty(m), np.empty(m)
    ni, pi = np.empty(m, np.int32), np.empty(m, np.int32)
    l, u = np.ascontiguousarray(l, np.float64), np.ascontiguousarray(u, np.float64)
    kr = np.ascontiguousarray(keep_row, np.uint8) if keep_row is not None else None
    kc = np.ascontiguousarray(keep_col, np.uint8) if keep_col is not None else None
    load().or_row_activity_summary(C.byref(view.lp.A), _p(l), _p(u), float(zero_tol),
                                   _p(kr) if kr is not None else None, _p(kc) if kc is not None else None,
                                   _p(mn), _p(mx), _p(ni), _p(pi))
    return mn, mx, ni, pi


class Plan:
    """Host plan + stats over a host LP for rule-level checks."""

    def __init__(self, lp: "G.HostLP", tape=True):
        self.view = HostLPView(lp)
        self.m, self.n = lp.m, lp.n
        self.plan = G.GpPlan()
        self.stats = G.GpStats()
        load().or_plan_create(C.byref(self.view.lp), C.byref(self.plan))
        load().or_stats_create(lp.m, lp.n, C.byref(self.stats))
        if tape:
            load().or_plan_enable_tape(C.byref(self.plan))

    def refresh(self):
        load().or_refresh_live_stats(C.byref(self.stats), C.byref(self.view.lp), C.byref(self.plan))

    def apply(self, rule, params):
        rid = G.RULE[rule] if isinstance(rule, str) else int(rule)
        return load().or_rule_apply(rid, C.byref(self.plan), C.byref(self.view.lp), C.byref(self.stats),
                                    C.byref(params))

    # host field access ------------------------------------------------------
    _FIELDS = {"keep_row_mask": ("m", np.uint8), "keep_col_mask": ("n", np.uint8), "new_c": ("n", np.float64),
               "new_l": ("n", np.float64), "new_u": ("n", np.float64), "new_AL": ("m", np.float64),
               "new_AU": ("m", np.float64)}
    _STATS = {"row_nnz": ("m", np.int32), "empty_row_mask": ("m", np.uint8), "singleton_row_mask": ("m", np.uint8),
              "singleton_row_col": ("m", np.int32), "singleton_row_val": ("m", np.float64),
              "col_nnz": ("n", np.int32), "empty_col_mask": ("n", np.uint8),
              "singleton_column_mask": ("n", np.uint8), "singleton_column_row": ("n", np.int32),
              "singleton_column_val": ("n", np.float64)}

    def get(self, name):
        if name in self._FIELDS:
            dim, dt = self._FIELDS[name]
            return _arr(getattr(self.plan, name), getattr(self, dim), dt)
        dim, dt = self._STATS[name]
        return _arr(getattr(self.stats, name), getattr(self, dim), dt)

    def set(self, name, values):
        src = self._FIELDS if name in self._FIELDS else self._STATS
        dim, dt = src[name]
        a = np.ascontiguousarray(values, dtype=dt)
        assert a.size == getattr(self, dim)
        ptr = getattr(self.plan if name in self._FIELDS else self.stats, name)
        if a.nbytes:
            C.memmove(ptr, a.ctypes.data, a.nbytes)

    def flags(self):
        return {k: int(getattr(self.plan, k)) for k in G.PLAN_FLAGS}

    def startup(self, rule, params, pending=True):
        """or_startup_rule into this plan; returns (applicable, count)."""
        rid = G.STARTUP_RULE[rule] if isinstance(rule, str) else int(rule)
        count = C.c_int32()
        rc = load().or_startup_rule(rid, C.byref(self.view.lp), C.byref(params), C.byref(self.plan),
                                    1 if pending else 0, C.byref(count))
        assert rc >= 0
        return bool(rc), count.value

    def tape(self):
        if not self.plan.tape:
            return None
        t = OrTape.from_address(self.plan.tape)
        return dict(types=_arr(t.types, t.count, np.int32), index_starts=_arr(t.istarts, t.count + 1, G.NP_OFFSET),
                    value_starts=_arr(t.vstarts, t.count + 1, G.NP_OFFSET),
                    dual_modes=_arr(t.modes, t.count, np.uint8), indices=_arr(t.idx, t.icount, np.int32),
                    vals=_arr(t.vals, t.vcount, np.float64))

    def free(self):
        if self.plan.keep_row_mask or self.plan.keep_col_mask:
            load().or_plan_destroy(C.byref(self.plan))
            load().or_stats_destroy(C.byref(self.stats))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- learned pass scheduler
def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Logits [batch][7] of the scheduler MLP (sequential FP32 restatement)."""
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, G.GP_MLP_IN)
    out = np.empty((X.shape[0], G.GP_MLP_OUT), dtype=np.float32)
    load().or_mlp_forward(_p(P), _p(X), X.shape[0], _p(out))
    return out


def mlp_loss_grad(params, x, y, mask):
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, G.GP_MLP_IN)
    Y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, G.GP_MLP_OUT)
    M = np.ascontiguousarray(mask, dtype=np.float32).reshape(-1, G.GP_MLP_OUT)
    grad = np.empty(G.GP_MLP_PARAMS, dtype=np.float32)
    loss = C.c_double(0.0)
    load().or_mlp_loss_grad(_p(P), _p(X), _p(Y), _p(M), X.shape[0], C.byref(loss), _p(grad))
    return float(loss.value), grad


def mlp_adam_step(params, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """In-place Adam step on float32 arrays."""
    load().or_mlp_adam_step(_p(params), _p(np.ascontiguousarray(grad, dtype=np.float32)), _p(m), _p(v), int(t),
                            float(lr), float(beta1), float(beta2), float(eps))


def schedule_features(trace, done, phase, nnz_now, nnz0, m0, n0, bound_only, limit) -> np.ndarray:
    arr = (G.GpRoundTrace * max(1, len(trace)))(*trace) if trace else (G.GpRoundTrace * 1)()
    f = np.empty(G.GP_MLP_IN, dtype=np.float32)
    load().or_schedule_features(arr, int(done), int(phase), int(nnz_now), int(nnz0), int(m0), int(n0),
                                int(bound_only), int(limit), _p(f))
    return f
