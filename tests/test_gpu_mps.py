"""MPS ingest on the GPU: gp_mps_upload (duplicates summed, A^T and A built
on the device) against the MPS restatement's CSC / CSR -- the reference's
known answers, seeded random texts, a 200k x 300k integer LP whose reading is
exact (compared with the host-CSR upload), entries in arbitrary order through
gp_lp_upload_entries, and the presolve of an ingested LP against the CPU
restatement's presolve of the same LP."""
import time

import numpy as np
import pytest

import gpupre as G
import mps_texts as T
import oracle as O
import oracle_mps as M
from test_mps_kats import check_expect, kat_cases

pytestmark = pytest.mark.gpu


def same_bits(x, y, what):
    x, y = np.asarray(x), np.asarray(y)
    assert x.shape == y.shape, f"{what}: shape {x.shape} vs {y.shape}"
    if x.dtype == np.float64:
        ok = (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))
    else:
        ok = x == y
    if not ok.all():
        k = int(np.flatnonzero(~ok)[0])
        raise AssertionError(f"{what}: {int((~ok).sum())} mismatches, first at {k}: {x[k]!r} vs {y[k]!r}")


def device_matches(d: G.DeviceLP, ref: M.MpsLp, what: str):
    h = G.device_lp_to_host(d.lp)
    assert (h.m, h.n) == (ref.m, ref.n), what
    same_bits(h.row_ptr.astype(np.int64), ref.row_ptr, f"{what} A row_ptr")
    same_bits(h.col_idx, ref.col_idx, f"{what} A col_idx")
    same_bits(h.vals, ref.csr_val, f"{what} A vals")
    same_bits(h.at[0].astype(np.int64), ref.col_ptr, f"{what} AT row_ptr")
    same_bits(h.at[1], ref.row_idx, f"{what} AT col_idx")
    same_bits(h.at[2], ref.csc_val, f"{what} AT vals")
    for k in ("c", "AL", "AU", "l", "u"):
        same_bits(getattr(h, k), getattr(ref, k), f"{what} {k}")
    same_bits(np.array([h.obj_constant]), np.array([ref.obj_constant]), f"{what} obj_constant")
    return h


@pytest.mark.parametrize("case", kat_cases(), ids=lambda c: c["test"])
def test_reference_known_answers_on_device(case):
    ref = M.read_mps(case["text"])
    with G.parse_mps(case["text"]) as model:
        with model.upload() as d:
            device_matches(d, ref, case["test"])
    check_expect(ref, case["expect"])


@pytest.mark.parametrize("block", range(2))
def test_random_texts_on_device(block):
    done = 0
    for seed in range(block * 150, block * 150 + 150):
        text = T.random_text(seed, rows=40, cols=60, defect_rate=0.0)
        try:
            ref = M.read_mps(text)
        except M.MpsReadError:
            continue  # a reading error is covered by the host tests
        print(f"[mps] seed {seed} {ref.m} x {ref.n} nnz={len(ref.csc_val)}", flush=True)
        with G.parse_mps(text) as model:
            with model.upload() as d:
                device_matches(d, ref, f"seed {seed}")
        done += 1
    assert done > 100


def test_integer_lp_200k_by_300k_reads_exactly():
    rng = np.random.default_rng(11)
    text, exp = T.write_integer_lp(rng, 200_000, 300_000, 2_000_000, dup_rate=0.05)
    print(f"[mps] integer LP text {len(text) / 1e6:.1f} MB", flush=True)
    t0 = time.perf_counter()
    model = G.parse_mps(text)
    t1 = time.perf_counter()
    try:
        d = model.upload()
        t2 = time.perf_counter()
    finally:
        model.free()
    host = G.HostLP(exp["m"], exp["n"], exp["row_ptr"], exp["col_idx"], exp["vals"], exp["c"], exp["l"], exp["u"],
                    exp["AL"], exp["AU"])
    print(f"[mps] parse {t1 - t0:.2f} s ({len(text) / (t1 - t0) / 1e6:.0f} MB/s), device assembly {t2 - t1:.3f} s, "
          f"{model.entries} entries -> {host.nnz} nonzeros", flush=True)
    with d, G.DeviceLP(host) as e:
        got, want = G.device_lp_to_host(d.lp), G.device_lp_to_host(e.lp)
    for k in ("row_ptr", "col_idx", "vals", "c", "AL", "AU", "l", "u"):
        same_bits(getattr(got, k), getattr(want, k), k)
    for a, b, k in zip(got.at, want.at, ("AT row_ptr", "AT col_idx", "AT vals")):
        same_bits(a, b, k)


def test_entries_in_any_order_with_duplicates():
    rng = np.random.default_rng(5)
    m, n, count = 3000, 2000, 60_000
    rows = rng.integers(0, m, count).astype(np.int32)
    cols = rng.integers(0, n, count).astype(np.int32)
    rows[1::7] = rows[0::7][: len(rows[1::7])]  # many repeated (row, column) pairs
    cols[1::7] = cols[0::7][: len(cols[1::7])]
    vals = rng.standard_normal(count)
    vals[::13] = 0.0  # explicit zeros are kept
    # expected: per (column, row) the values in input order, summed from 0.0
    sums = {}
    for i, j, v in zip(rows.tolist(), cols.tolist(), vals.tolist()):
        sums[(j, i)] = sums.get((j, i), 0.0) + v
    keys = sorted(sums)
    col_ptr = np.searchsorted(np.array([k[0] for k in keys]), np.arange(n + 1))
    vec = [rng.standard_normal(n), rng.standard_normal(m), rng.standard_normal(m) + 5, -np.ones(n), np.ones(n)]
    with G.DeviceLP.from_entries(m, n, rows, cols, vals, *vec, obj_constant=2.5) as d:
        h = G.device_lp_to_host(d.lp)
    same_bits(h.at[0].astype(np.int64), col_ptr.astype(np.int64), "AT row_ptr")
    same_bits(h.at[1], np.array([k[1] for k in keys], np.int32), "AT col_idx")
    same_bits(h.at[2], np.array([sums[k] for k in keys]), "AT vals")
    csr = sorted((k[1], k[0]) for k in keys)
    same_bits(h.col_idx, np.array([k[1] for k in csr], np.int32), "A col_idx")
    same_bits(h.vals, np.array([sums[(k[1], k[0])] for k in csr]), "A vals")
    assert h.obj_constant == 2.5


def test_entry_outside_the_matrix_is_rejected():
    rows, cols = np.array([0, 3], np.int32), np.array([0, 0], np.int32)
    with pytest.raises(G.GpError) as e:
        G.DeviceLP.from_entries(3, 2, rows, cols, np.ones(2), np.zeros(2), np.zeros(3), np.zeros(3),
                                np.zeros(2), np.ones(2))
    assert e.value.code == 1 and "outside" in str(e.value)


@pytest.mark.parametrize("seed", [0, 1])
def test_presolve_of_ingested_lp_matches_restatement(seed):
    rng = np.random.default_rng(100 + seed)
    text, _ = T.write_integer_lp(rng, 3000, 4500, 20_000)
    ref = M.read_mps(text)
    host = G.HostLP(ref.m, ref.n, ref.row_ptr.astype(np.int32), ref.col_idx, ref.csr_val, ref.c, ref.l, ref.u,
                    ref.AL, ref.AU, ref.obj_constant)
    params = G.parity_params()
    ores = O.presolve(host, params)
    with G.parse_mps(text) as model, model.upload() as d:
        s = G.preprocess(d, params)
        try:
            print(f"[mps] presolve {ref.m} x {ref.n} -> {s.reduced_rows} x {s.reduced_cols}", flush=True)
            assert (s.reduced_rows, s.reduced_cols, s.iterations) == (ores.m1, ores.n1, ores.iterations)
            assert s.has_infeasible == ores.has_infeasible
            red, rec = s.reduced(), s.record()
            for k in ("row_ptr", "col_idx", "vals", "c", "l", "u", "AL", "AU"):
                same_bits(getattr(red, k), getattr(ores.reduced, k), f"reduced {k}")
            same_bits(rec.row_red2org, ores.row_red2org, "row_red2org")
            same_bits(rec.col_red2org, ores.col_red2org, "col_red2org")
            for k in ("types", "indices", "vals"):
                same_bits(getattr(rec, k), getattr(ores, k), f"tape {k}")
        finally:
            s.free()
