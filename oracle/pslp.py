"""CPU baseline only: ctypes front of _ref/libpslp.so, the reference's bundled
CPU presolver PSLP 0.0.11 built from its sources by ref.mk."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libpslp.so")
_lib = None


def available() -> bool:
    return os.path.exists(PATH)


def _load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(PATH)
        _lib.pslp_presolve_timed.restype = C.c_int
        _lib.pslp_presolve_timed.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 3 + [C.c_void_p] * 6
    return _lib


def presolve(lp) -> dict:
    """Run PSLP with default settings on a gpupre.HostLP; returns wall seconds
    (new_presolver + run_presolver), PSLP's own time_presolve, status and the
    reduced dimensions."""
    arrs = [np.ascontiguousarray(a, dtype=t) for a, t in
            ((lp.vals, np.float64), (lp.col_idx, np.int32), (lp.row_ptr, np.int32), (lp.AL, np.float64),
             (lp.AU, np.float64), (lp.l, np.float64), (lp.u, np.float64), (lp.c, np.float64))]
    Ax, Ai, Ap, lhs, rhs, lb, ub, c = arrs
    out = np.zeros(5)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    status = _load().pslp_presolve_timed(p(Ax), p(Ai), p(Ap), lp.m, lp.n, int(lp.nnz), p(lhs), p(rhs), p(lb), p(ub),
                                         p(c), p(out))
    return {"status": int(status), "wall_seconds": float(out[0]), "time_presolve": float(out[1]),
            "rows": int(out[2]), "cols": int(out[3]), "nnz": int(out[4])}
