"""The premise of the batched presolve, on the CPU restatement: presolving the
block-diagonal direct sum of several LPs never couples two blocks.  Every
reduced row keeps only columns of its own block, every tape record refers to
one block, and the reduced blocks keep the block order -- so the direct sum's
presolve splits into one presolve per instance (gp_preprocess_batch)."""
import numpy as np

import gpupre as G
import oracle as O


def instances():
    out = [G.generate(m, m * 3 // 2, seed=40 + k, degree_mode=0, mean_degree=6.0)
           for k, m in enumerate((1000, 2500, 1200, 4000))]
    out.insert(2, G.generate_qap(4, seed=9))
    return out


def block_of(offsets, idx):
    return np.searchsorted(offsets, idx, side="right") - 1


def test_direct_sum_presolve_keeps_blocks():
    lps = instances()
    big = G.direct_sum(lps)
    ro = np.cumsum([0] + [lp.m for lp in lps])
    co = np.cumsum([0] + [lp.n for lp in lps])
    assert big.m == ro[-1] and big.n == co[-1] and big.nnz == sum(lp.nnz for lp in lps)
    res = O.presolve(big, G.parity_params())
    assert not res.has_infeasible and not res.has_unbounded
    rb = block_of(ro, res.row_red2org)
    cb = block_of(co, res.col_red2org)
    assert np.all(np.diff(rb) >= 0) and np.all(np.diff(cb) >= 0)  # block order kept
    red = res.reduced
    rows = np.repeat(np.arange(red.m), np.diff(red.row_ptr))
    assert np.array_equal(rb[rows], cb[red.col_idx])  # no entry couples two blocks
    # every tape record's indices lie in one block (rows and columns of the
    # original direct sum; a record mixes row and column indices, so accept a
    # record when all its indices fit one block's rows-or-columns ranges)
    for r in range(res.types.size):
        idx = res.indices[res.index_starts[r]:res.index_starts[r + 1]]
        if res.types[r] == 13:  # DoubletonEquation {elim, keep, row, 0}
            idx = idx[:3]
        elif res.types[r] == 2:  # SingletonColumnSubstitution [col, row, support_count, support...]
            idx = np.concatenate([idx[:2], idx[3:]])
        idx = idx[idx >= 0]
        if idx.size == 0:
            continue
        ok = False
        for k in range(len(lps)):
            in_rows = (idx >= ro[k]) & (idx < ro[k + 1])
            in_cols = (idx >= co[k]) & (idx < co[k + 1])
            if np.all(in_rows | in_cols):
                ok = True
                break
        assert ok, (r, res.types[r], idx)


def test_single_instance_direct_sum_is_the_instance():
    lp = G.generate(3000, 4500, seed=77, degree_mode=0, mean_degree=6.0)
    big = G.direct_sum([lp])
    for a, b in ((big.row_ptr, lp.row_ptr), (big.col_idx, lp.col_idx), (big.vals, lp.vals), (big.c, lp.c),
                 (big.AL, lp.AL), (big.AU, lp.AU), (big.l, lp.l), (big.u, lp.u)):
        assert np.array_equal(a, b)
