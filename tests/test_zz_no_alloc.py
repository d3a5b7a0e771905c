"""gp_row_activity_summary_noalloc (scratch in a fixed device array, no
allocation) against gp_row_activity_summary (pool scratch) through the C ABI:
bit-identical outputs on two calls in a row, the second on a smaller LP, and
a clear error beyond the scratch limit."""
import ctypes as C

import numpy as np
import pytest

import gpupre as G
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

MAX_COLS = 1 << 21


def summaries(fn, d, l, u, keep_row, keep_col, zero_tol):
    m = d.m
    outs = [G.DeviceArray(8 * m), G.DeviceArray(8 * m), G.DeviceArray(4 * m), G.DeviceArray(4 * m)]
    dl, du = G.DeviceArray.from_host(l), G.DeviceArray.from_host(u)
    kr = G.DeviceArray.from_host(keep_row)
    kc = G.DeviceArray.from_host(keep_col) if keep_col is not None else None
    G.check(fn(C.byref(d.lp.A), dl.ptr, du.ptr, C.c_double(zero_tol), kr.ptr, kc.ptr if kc else None,
               *[o.ptr for o in outs]))
    return (G.to_host(outs[0].ptr, m, np.float64), G.to_host(outs[1].ptr, m, np.float64),
            G.to_host(outs[2].ptr, m, np.int32), G.to_host(outs[3].ptr, m, np.int32))


def test_noalloc_activity_summary_matches_pooled_entry_point():
    lib = G.load()
    cases = [("medium", 0, True), ("small", 3, False)]  # second call: smaller LP, no column mask
    for name, seed, masked in cases:
        print(f"[no-alloc] {name} seed={seed} masked={masked}", flush=True)
        lp = G.generate_workload(name, seed=seed)
        assert lp.n <= MAX_COLS
        rng = np.random.default_rng(seed + 11)
        l, u = lp.l.copy(), lp.u.copy()
        u[rng.random(lp.n) < 0.05] = np.inf
        l[rng.random(lp.n) < 0.05] = -np.inf
        keep_row = (rng.random(lp.m) > 0.1).astype(np.uint8)
        keep_col = (rng.random(lp.n) > 0.1).astype(np.uint8) if masked else None
        with G.DeviceLP(lp) as d:
            ref = summaries(lib.gp_row_activity_summary, d, l, u, keep_row, keep_col, 1e-9)
            got = summaries(lib.gp_row_activity_summary_noalloc, d, l, u, keep_row, keep_col, 1e-9)
        for a, b, k in zip(got, ref, ("min_fin", "max_fin", "min_ninf", "max_pinf")):
            assert_same(a, b, f"{name}: {k}")


def test_noalloc_rejects_beyond_scratch_limit():
    lib = G.load()
    big = G.GpCsr(1, MAX_COLS + 1, 0, None, None, None)
    dummy = G.DeviceArray(64)
    rc = lib.gp_row_activity_summary_noalloc(C.byref(big), dummy.ptr, dummy.ptr, C.c_double(0.0), dummy.ptr, None,
                                             dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr)
    assert rc == 1  # GP_ERR_INVALID_ARGUMENT, before any device work
    assert b"2097152" in lib.gp_last_error()
