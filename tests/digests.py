"""SHA-256 digests of a presolve result (restatement or engine), so results
of the full-size workloads can be compared without keeping either side."""
import hashlib

import numpy as np


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest(iterations, flags, m1, n1, obj_delta, trace, rec, red):
    """flags = (has_infeasible, has_unbounded, has_change); rec has the record
    arrays (row_red2org ... vals), red the reduced LP (row_ptr ... AU)."""
    return {
        "iterations": int(iterations), "has_infeasible": int(flags[0]), "has_unbounded": int(flags[1]),
        "has_change": int(flags[2]), "m1": int(m1), "n1": int(n1), "nnz1": int(len(red.col_idx)),
        "obj_constant_delta_bits": int(np.float64(obj_delta).view(np.uint64)),
        "trace": [[int(t.phase), int(t.next_phase), int(t.nnz_before), int(t.nnz_after), int(t.fired_mask)]
                  for t in trace],
        "sha256": {
            "row_red2org": sha(rec.row_red2org), "col_red2org": sha(rec.col_red2org),
            "fixed_idx": sha(rec.fixed_idx), "fixed_val": sha(rec.fixed_val),
            "types": sha(rec.types), "index_starts": sha(rec.index_starts), "value_starts": sha(rec.value_starts),
            "dual_modes": sha(rec.dual_modes), "indices": sha(rec.indices), "vals": sha(rec.vals),
            "row_ptr": sha(red.row_ptr), "col_idx": sha(red.col_idx), "a_vals": sha(red.vals), "c": sha(red.c),
            "l": sha(red.l), "u": sha(red.u), "AL": sha(red.AL), "AU": sha(red.AU),
        },
    }


def digest_oracle(res):
    return digest(res.iterations, (res.has_infeasible, res.has_unbounded, res.has_change), res.m1, res.n1,
                  res.obj_constant_delta, res.trace, res, res.reduced)


def digest_engine(summary):
    return digest(summary.iterations, (summary.has_infeasible, summary.has_unbounded, summary.has_change),
                  summary.reduced_rows, summary.reduced_cols, summary.obj_constant_delta, summary.trace(),
                  summary.record(), summary.reduced())


def compare(got, exp, what):
    for k in ("iterations", "has_infeasible", "has_unbounded", "has_change", "m1", "n1", "nnz1",
              "obj_constant_delta_bits"):
        assert got[k] == exp[k], f"{what}: {k} {got[k]} vs {exp[k]}"
    assert got["trace"] == exp["trace"], f"{what}: schedule trace differs"
    bad = [k for k in exp["sha256"] if got["sha256"][k] != exp["sha256"][k]]
    assert not bad, f"{what}: arrays differ: {bad}"
