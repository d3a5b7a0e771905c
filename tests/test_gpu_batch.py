"""Batched presolve on the GPU (gp_preprocess_batch).

* assembly: with presolve off the batch's direct sum is bit-identical to the
  host direct sum of the instances;
* parity: the batch equals the CPU restatement's presolve of the host direct
  sum bit for bit (maps, tape, reduced LP, objective constant);
* split: every extracted reduced instance, solved by HiGHS, plus its
  per-instance objective constant, reaches the instance's own optimum within
  1e-6 (relative) -- which pins the per-instance objective constants;
* postsolve: the whole reduced batch solved by HiGHS and postsolved returns
  per-instance points with the optimal objective and primal feasibility;
* a batch of one instance is the plain presolve of that instance.
"""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from lp_check import primal_residual, relative_gap, solve

pytestmark = pytest.mark.gpu


def instances():
    out = [G.generate(m, m * 3 // 2, seed=40 + k, degree_mode=0, mean_degree=6.0)
           for k, m in enumerate((1000, 2500, 1200, 3000, 1500))]
    out.insert(2, G.generate_qap(4, seed=9))
    out.append(G.HostLP.from_dense(np.array([[1.0, 2.0]]), np.array([1.0, 1.0]), np.array([0.0, 0.0]),
                                   np.array([4.0, 4.0]), np.array([1.0]), np.array([3.0]), obj_constant=0.5))
    return out


def run_batch(lps, params=None):
    dl = [G.DeviceLP(lp) for lp in lps]
    b = G.preprocess_batch(dl, params or G.parity_params())
    return dl, b


def free_all(dl, b):
    b.free()
    for d in dl:
        d.free()


def test_assembly_matches_host_direct_sum():
    lps = instances()
    dl, b = run_batch(lps, G.parity_params(enable_presolve=0))
    try:
        got = b.whole().reduced()
        ref = G.direct_sum(lps)
        for name in ("row_ptr", "col_idx", "c", "l", "u", "AL", "AU"):
            assert np.array_equal(getattr(got, name), getattr(ref, name)), name
        assert np.array_equal(got.vals.view(np.uint64), ref.vals.view(np.uint64))
        atp, ati, atv = ref.transpose()
        assert np.array_equal(got.at[0], atp) and np.array_equal(got.at[1], ati)
        assert np.array_equal(got.at[2].view(np.uint64), atv.view(np.uint64))
        assert got.obj_constant == ref.obj_constant
        assert np.array_equal(b.offsets("row_offsets"), np.cumsum([0] + [lp.m for lp in lps]))
    finally:
        free_all(dl, b)


def test_batch_matches_restatement_of_direct_sum():
    lps = instances()
    dl, b = run_batch(lps)
    try:
        w = b.whole()
        rec = w.record()
        ref = O.presolve(G.direct_sum(lps), G.parity_params())
        assert (w.reduced_rows, w.reduced_cols, w.reduced_nnz) == (ref.m1, ref.n1, ref.reduced.nnz)
        assert np.array_equal(rec.row_red2org, ref.row_red2org) and np.array_equal(rec.col_red2org, ref.col_red2org)
        assert np.array_equal(rec.types, ref.types) and np.array_equal(rec.indices, ref.indices)
        assert np.array_equal(rec.vals.view(np.uint64), ref.vals.view(np.uint64))
        red = w.reduced()
        assert np.array_equal(red.vals.view(np.uint64), ref.reduced.vals.view(np.uint64))
        assert w.obj_constant_delta == ref.obj_constant_delta
        # reduced offsets agree with the maps
        ro = np.cumsum([0] + [lp.m for lp in lps])
        co = np.cumsum([0] + [lp.n for lp in lps])
        assert np.array_equal(b.offsets("reduced_row_offsets"), np.searchsorted(ref.row_red2org, ro))
        assert np.array_equal(b.offsets("reduced_col_offsets"), np.searchsorted(ref.col_red2org, co))
        objs = b.obj_constants()
        total = sum(lp.obj_constant for lp in lps) + ref.obj_constant_delta
        assert abs(objs.sum() - total) <= 1e-9 * max(1.0, abs(total))
    finally:
        free_all(dl, b)


def test_batch_instances_reach_their_optima():
    lps = instances()
    dl, b = run_batch(lps)
    try:
        objs = b.obj_constants()
        for k, lp in enumerate(lps):
            st0, obj0, *_ = solve(lp)
            assert st0 == 0
            red = b.instance(k)
            assert red.obj_constant == objs[k]
            if red.n == 0:
                got = red.obj_constant
            else:
                st, got, *_ = solve(red)
                assert st == 0
            assert relative_gap(got, obj0) <= 1e-6, (k, got, obj0)
    finally:
        free_all(dl, b)


def test_batch_postsolve():
    lps = instances()
    dl, b = run_batch(lps)
    try:
        red = b.whole().reduced()
        st, _, xr, yr, zr = solve(red)
        assert st == 0
        x, _, _ = b.postprocess(xr, yr, zr)
        co = np.cumsum([0] + [lp.n for lp in lps])
        for k, lp in enumerate(lps):
            xk = x[co[k]:co[k + 1]]
            _, obj0, *_ = solve(lp)
            assert relative_gap(float(lp.c @ xk) + lp.obj_constant, obj0) <= 1e-6
            assert primal_residual(lp, xk) <= 1e-6
    finally:
        free_all(dl, b)


def test_batch_of_one_is_plain_presolve():
    lp = G.generate(3000, 4500, seed=77, degree_mode=0, mean_degree=6.0)
    dl, b = run_batch([lp])
    try:
        with G.DeviceLP(lp) as d:
            s = G.preprocess(d, G.parity_params())
            try:
                r1, r2 = s.record(), b.whole().record()
                assert np.array_equal(r1.types, r2.types) and np.array_equal(r1.indices, r2.indices)
                assert np.array_equal(r1.vals.view(np.uint64), r2.vals.view(np.uint64))
                assert s.obj_constant_delta == b.whole().obj_constant_delta
                # per-instance sums run in another fixed order than the global one
                want = lp.obj_constant + s.obj_constant_delta
                assert abs(b.obj_constants()[0] - want) <= 1e-12 * max(1.0, abs(want))
            finally:
                s.free()
    finally:
        free_all(dl, b)


def test_batch_rejects_bad_input():
    with pytest.raises(G.GpError):
        G.check(G.load().gp_preprocess_batch(None, 0, G.parity_params(), None))
