"""Per-instance termination of the batched presolve: an instance that a rule
finds infeasible or unbounded retires alone.

The batch then equals the CPU restatement's presolve of the direct sum of the
other instances, bit for bit, with the retired instances as empty blocks and
their status reported; without such instances nothing changes."""
import numpy as np
import pytest

import gpupre as G
import oracle as O
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def with_empty_row(lp, side):
    """lp plus one row without entries whose sides are [side, side]."""
    rp = np.concatenate([lp.row_ptr, lp.row_ptr[-1:]])
    return G.HostLP(lp.m + 1, lp.n, rp, lp.col_idx, lp.vals, lp.c, lp.l, lp.u,
                    np.concatenate([lp.AL, [side]]), np.concatenate([lp.AU, [side]]), lp.obj_constant)


def with_empty_column(lp, cost, upper):
    """lp plus one column without entries: cost, bounds [0, upper]."""
    return G.HostLP(lp.m, lp.n + 1, lp.row_ptr, lp.col_idx, lp.vals, np.concatenate([lp.c, [cost]]),
                    np.concatenate([lp.l, [0.0]]), np.concatenate([lp.u, [upper]]), lp.AL, lp.AU, lp.obj_constant)


def instances():
    return [G.generate(m, m * 3 // 2, seed=70 + k, degree_mode=0, mean_degree=6.0)
            for k, m in enumerate((1000, 1800, 1200, 2600, 1500, 2200, 1100, 1300))]


def test_infeasible_and_unbounded_instances_retire_alone():
    lps = instances()
    lps[2] = with_empty_row(lps[2], 1.0)               # 0 = 1: infeasible
    lps[5] = with_empty_column(lps[5], -1.0, np.inf)  # cost -1, no upper bound: unbounded
    print(f"[batch retire] {len(lps)} instances, planted infeasible 2, unbounded 5", flush=True)
    dl = [G.DeviceLP(lp) for lp in lps]
    b = G.preprocess_batch(dl, G.parity_params())
    try:
        status = b.status()
        assert list(status) == [0, 0, 1, 0, 0, 2, 0, 0], status
        assert b.restarts >= 1
        w = b.whole()
        assert not w.has_infeasible and not w.has_unbounded
        keep = [lp for k, lp in enumerate(lps) if status[k] == 0]
        ref = O.presolve(G.direct_sum(keep), G.parity_params())
        rec, red = w.record(), w.reduced()
        assert (w.reduced_rows, w.reduced_cols, w.reduced_nnz) == (ref.m1, ref.n1, ref.reduced.nnz)
        assert_same(rec.row_red2org, ref.row_red2org, "row_red2org")
        assert_same(rec.col_red2org, ref.col_red2org, "col_red2org")
        for k in ("types", "index_starts", "value_starts", "dual_modes", "indices", "vals"):
            assert_same(getattr(rec, k), getattr(ref, k), f"tape {k}")
        for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
            assert_same(getattr(red, k), getattr(ref.reduced, k), f"reduced {k}")
        # the retired instances are empty blocks of the presolved direct sum
        ro, co = b.offsets("row_offsets"), b.offsets("col_offsets")
        rro, rco = b.offsets("reduced_row_offsets"), b.offsets("reduced_col_offsets")
        for k in (2, 5):
            assert ro[k + 1] == ro[k] and co[k + 1] == co[k]
            assert rro[k + 1] == rro[k] and rco[k + 1] == rco[k]
        assert ro[-1] == sum(lp.m for lp in keep) and co[-1] == sum(lp.n for lp in keep)
        # each retired instance is infeasible / unbounded alone, too
        assert O.presolve(lps[2], G.parity_params()).has_infeasible
        assert O.presolve(lps[5], G.parity_params()).has_unbounded
    finally:
        b.free()
        for d in dl:
            d.free()


def test_batch_without_terminal_instances_is_unchanged():
    lps = instances()
    dl = [G.DeviceLP(lp) for lp in lps]
    b = G.preprocess_batch(dl, G.parity_params())
    try:
        assert list(b.status()) == [0] * len(lps) and b.restarts == 0
        ref = O.presolve(G.direct_sum(lps), G.parity_params())
        w = b.whole()
        assert (w.reduced_rows, w.reduced_cols, w.reduced_nnz) == (ref.m1, ref.n1, ref.reduced.nnz)
        assert_same(w.record().types, ref.types, "tape types")
    finally:
        b.free()
        for d in dl:
            d.free()
