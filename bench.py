#!/usr/bin/env python3
"""Benchmark: full LP presolve of the 5M x 8M (~61M nnz, Zipf column degree)
synthetic LP on one MI355X.

A step is one complete presolve (gp_preprocess: every round to the fixpoint,
with the recovery tape) of the LP already resident in HBM, followed by
gp_free_summary.  Prints one JSON line.  `roofline` is measured on the
dominant kernel family with HIP events (algorithmic bytes per launch divided
by the average launch time).  `cpu_baseline` times the reference's bundled CPU
presolver (PSLP 0.0.11, built from its sources into oracle/_ref) on the same
full LP; `cpu_restatement` times this engine's CPU restatement on a bounded
sample of the same recipe.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-preprocessor-ijoc_amd"))

HBM_PEAK_GBPS = 8000.0

# kernel-timing family -> the one kernel it brackets (rocprofv3 kernel name)
KERNEL_OF_FAMILY = {
    "compact": "k_compact_tiles",
    "activity": "k_row_fold<ActivityFold, 256, 2048>",
    "sc_activity": "k_row_fold<ScActivityFold, 256, 2048>",
    "redundant": "k_row_fold<RedundantFold, 256, 2048>",
    "bound_candidates": "k_bs_candidates",
    "col_max": "k_seg_reduce_tiles<AbsMaxOp>",
}

# what each bench workload is (the generator recipe in csrc/lpgen.c)
RECIPE = ("Zipf(1.5) column degree capped at 500, 1% empty columns; planted 5% singleton rows, 5% doubleton "
          "equalities, 2% parallel rows, 2% duplicate and 3% fixed columns; 40% equality rows; 10% free columns; "
          "int32 indices, FP64 values")
WORKLOAD_TEXT = {
    "large": "5M x 8M synthetic LP, " + RECIPE + "; integer-valued data: coefficients +-1..4, integer bounds, "
             "planted point and slacks, parallel-row factors in {+-1, +-2, +-3}, so every activity and shift is "
             "exact in binary64",
    "large_fp64": "5M x 8M synthetic LP, " + RECIPE + "; coefficients N(0,1) clipped away from |a| < 1e-3, "
                  "parallel-row factors U[0.5, 4]",
    "large10": "500k x 800k, the large recipe at 1/10 scale, integer-valued data",
    "medium": "200k x 300k synthetic LP, ~2M nnz, " + RECIPE,
    "small": "2k x 3k synthetic LP, uniform degree 6, FP64 N(0,1) coefficients",
}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by the copy-kernel calibration),
    or None when no such measurement is committed."""
    path = os.path.join(REPO, "profiles", f"pmc_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    row = table.get("kernels", {}).get(kernel)
    return None if row is None else row["bytes_per_launch"]
METRIC = "presolve ms + HBM GB/s/round; batched LPs/s; learned vs plain at 1e-6 gap"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(v, q):
    import numpy as np

    return round(float(np.percentile(np.asarray(v) * 1e3, q)), 3)


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


def pslp_pinned(lp, cores=2):
    """PSLP 0.0.11 on `lp`, its threads pinned to the first `cores` CPUs this
    process may use (taskset -c 0-1 in the reference's setup); PSLP's prints
    go to stderr so stdout keeps the one JSON line."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pslp

    if not pslp.available():
        return None
    saved_aff = os.sched_getaffinity(0)
    use = set(sorted(saved_aff)[:cores])
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        os.sched_setaffinity(0, use)
        r = pslp.presolve(lp)
    finally:
        os.sched_setaffinity(0, saved_aff)
        import ctypes

        ctypes.CDLL(None).fflush(None)  # PSLP's buffered C stdio goes out while fd 1 is still stderr
        os.dup2(saved, 1)
        os.close(saved)
    st = r["status"]
    r["outcome"] = ("infeasible" if st >= 0 and st & 2 else "unbounded" if st >= 0 and st & 4 else
                    "reduced" if st >= 0 and st & 1 else "unchanged" if st == 0 else "error")
    r["cores"] = len(use)
    return r


def gpu_presolve_ms(lp, params, reps):
    """p50 and p99 wall time of gp_preprocess + gp_free_summary on a resident LP."""
    import gpupre as G

    with G.DeviceLP(lp) as d:
        G.preprocess(d, params).free()
        times, red = [], None
        for _ in range(reps):
            G.check(G.load().gp_synchronize())
            t = time.perf_counter()
            s = G.preprocess(d, params)
            red = (s.reduced_rows, s.reduced_cols, s.reduced_nnz, s.iterations,
                   "infeasible" if s.has_infeasible else "unbounded" if s.has_unbounded else "ok")
            s.free()
            times.append(time.perf_counter() - t)
    return pct(times, 50), pct(times, 99), red


def configs_leg(seed):
    """Presolve time of the benchmark configurations that fit on the host
    baseline as well: GPU p50 over 10 runs against PSLP on two pinned cores."""
    import gpupre as G

    cases = [("config1 small 2000x3000", G.generate_workload("small", seed)),
             ("config2 medium 200000x300000", G.generate_workload("medium", seed)),
             ("config5 QAP n=12", G.generate_qap(12, seed)),
             ("config5 QAP n=24", G.generate_qap(24, seed)),
             ("config5 generator 200000x300000", G.generate(200_000, 300_000, seed, degree_mode=0, mean_degree=6.0))]
    out = {}
    for name, lp in cases:
        p50, p99, red = gpu_presolve_ms(lp, G.parity_params(), 10)
        row = {"rows": lp.m, "cols": lp.n, "nnz": lp.nnz, "gpu_ms_p50": p50, "gpu_ms_p99": p99,
               "reduced": {"rows": red[0], "cols": red[1], "nnz": red[2], "iterations": red[3], "status": red[4]}}
        r = pslp_pinned(lp)
        if r is not None:
            row["pslp"] = {"time_presolve_ms": round(r["time_presolve"] * 1e3, 3),
                           "wall_ms": round(r["wall_seconds"] * 1e3, 3), "outcome": r["outcome"],
                           "reduced": [r["rows"], r["cols"], r["nnz"]], "cores": r["cores"]}
        out[name] = row
        log(f"{name}: {json.dumps(row)}")
    return out


def fp64_verdict_leg(seed):
    """The large recipe with FP64 N(0,1) data: time for the presolve to reach
    its verdict (the tolerance policy ends it as infeasible)."""
    import gpupre as G

    lp = G.generate_workload("large_fp64", seed)
    p50, p99, red = gpu_presolve_ms(lp, G.parity_params(), 3)
    out = {"workload": WORKLOAD_TEXT["large_fp64"], "rows": lp.m, "cols": lp.n, "nnz": lp.nnz, "ms_p50": p50,
           "iterations": red[3], "status": red[4], "reduced": {"rows": red[0], "cols": red[1], "nnz": red[2]}}
    log(f"large_fp64 verdict: {json.dumps(out)}")
    return out


def batched_leg(count, seed, serial=True):
    """Batched presolve of `count` LPs (rows log-uniform in [1e3, 1e5], columns
    1.5 x rows, uniform degree 6, FP64 N(0,1) coefficients) resident in HBM:
    one gp_preprocess_batch against the plain path presolving the same LPs one
    after another.  An instance a rule finds infeasible or unbounded retires
    alone (reported under "retired")."""
    import numpy as np
    import gpupre as G

    t = time.time()
    lps = G.batch_workload(count, seed, integral=0)
    dl = [G.DeviceLP(lp) for lp in lps]
    nnz = sum(lp.nnz for lp in lps)
    log(f"batched workload: {count} LPs, {sum(lp.m for lp in lps)} rows, {nnz} nnz "
        f"(generated + uploaded in {time.time() - t:.1f}s)")
    params = G.parity_params()
    # launch schedule tuned on the direct sum the batch presolves (without the
    # instances it retires), which the first batch run assembles
    t = time.time()
    b0 = G.preprocess_batch(dl, params)
    tune, tuned_ms, default_ms = G.autotune(b0.s.assembled, params, reps=1)
    b0.free()
    G.set_tuning(tune)
    log(f"autotuned schedule {tune.as_dict()}: presolve {tuned_ms:.1f} ms vs {default_ms:.1f} ms untuned "
        f"({time.time() - t:.1f}s)")
    G.preprocess_batch(dl, params).free()  # warm-up (pool)
    walls = []
    for _ in range(3):
        G.check(G.load().gp_synchronize())
        t = time.perf_counter()
        b = G.preprocess_batch(dl, params)
        walls.append(time.perf_counter() - t)
        w = b.whole()
        red = (w.reduced_rows, w.reduced_cols, w.reduced_nnz, w.iterations,
               "infeasible" if w.has_infeasible else "unbounded" if w.has_unbounded else "ok")
        asm, pre = b.s.assemble_seconds, b.s.presolve_seconds
        st = b.status()
        retired = {int(k): ("infeasible" if st[k] == 1 else "unbounded") for k in np.nonzero(st)[0]}
        restarts = b.restarts
        b.free()
    batch_s = min(walls)
    G.set_tuning()
    log(f"batched: {batch_s * 1e3:.1f} ms (assembly {asm * 1e3:.1f} ms, presolve {pre * 1e3:.1f} ms) -> {red}")
    lat = []
    G.check(G.load().gp_synchronize())
    t0 = time.perf_counter()
    for d in (dl if serial else []):
        t = time.perf_counter()
        G.preprocess(d, params).free()
        lat.append(time.perf_counter() - t)
    serial_s = time.perf_counter() - t0
    if serial:
        log(f"plain, one LP after another: {serial_s * 1e3:.1f} ms, p50 {pct(lat, 50)} ms, p99 {pct(lat, 99)} ms")
    else:
        lat, serial_s = [float("nan")], float("nan")
    for d in dl:
        d.free()
    return {"lps": count, "nnz": nnz, "lps_per_s": round(count / batch_s, 2), "ms": round(batch_s * 1e3, 2),
            "assembly_ms": round(asm * 1e3, 2), "latency_ms_p50": round(batch_s * 1e3, 2),
            "latency_ms_p99": round(batch_s * 1e3, 2),
            "reduced": {"rows": red[0], "cols": red[1], "nnz": red[2], "iterations": red[3], "status": red[4]},
            "retired": retired, "restarts": restarts,
            "plain_serial": {"lps_per_s": round(count / serial_s, 2), "ms": round(serial_s * 1e3, 2),
                             "latency_ms_p50": pct(lat, 50), "latency_ms_p99": pct(lat, 99)} if serial else None,
            "speedup_vs_plain": round(serial_s / batch_s, 2) if serial else None,
            "tuning": tune.as_dict(), "tuned_presolve_ms": round(tuned_ms, 2),
            "untuned_presolve_ms": round(default_ms, 2),
            "note": "every instance of a batch completes with the batch, so its per-instance latency is the "
                    "batch wall time"}


def learned_leg(seed):
    """Learned pass scheduler against the plain schedule on held-out LPs of both
    training families (QAP n in {14, 18, 22}; the synthetic generator at 200k
    rows, 3 seeds): best-of-3 presolve wall time and rule launches."""
    import gpupre as G
    import learned as L

    if not os.path.exists(L.WEIGHTS_PATH):
        return None
    blob = L.load_blob()
    cases = [("qap", G.generate_qap(n, seed=1_000_000 + seed + n)) for n in (14, 18, 22)]
    cases += [("gen200k", G.generate(200_000, 300_000, seed=1_000_001 + seed + k, degree_mode=0, mean_degree=6.0))
              for k in range(3)]
    out = {}
    for fam, lp in cases:
        with G.DeviceLP(lp) as d:
            for name, p in (("plain", G.parity_params(collect_trace=1)), ("learned", L.learned_params(blob, collect_trace=1))):
                G.preprocess(d, p).free()
                best, launches, nnz = None, 0, 0
                for _ in range(3):
                    G.check(G.load().gp_synchronize())
                    t = time.perf_counter()
                    s = G.preprocess(d, p)
                    dt = time.perf_counter() - t
                    launches, nnz = L.rule_launches(s.trace()), s.reduced_nnz
                    s.free()
                    best = dt if best is None else min(best, dt)
                row = out.setdefault(fam, {}).setdefault(name, {"ms": 0.0, "rule_launches": 0, "reduced_nnz": 0})
                row["ms"] += best * 1e3
                row["rule_launches"] += launches
                row["reduced_nnz"] += nnz
    for fam in out:
        for name in out[fam]:
            out[fam][name]["ms"] = round(out[fam][name]["ms"], 3)
        out[fam]["speedup"] = round(out[fam]["plain"]["ms"] / out[fam]["learned"]["ms"], 3)
    log(f"learned vs plain: {json.dumps(out)}")
    return out


def highs_objective(lp):
    """FP64 CPU reference solve (HiGHS through SciPy) of min c'x + k, AL <= Ax <= AU, l <= x <= u."""
    import numpy as np
    import scipy.sparse as sp
    from scipy.optimize import linprog

    A = sp.csr_matrix((lp.vals, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
    AL, AU = np.asarray(lp.AL), np.asarray(lp.AU)
    eq = np.isfinite(AL) & np.isfinite(AU) & (AL == AU)
    up, lo = np.isfinite(AU) & ~eq, np.isfinite(AL) & ~eq
    A_ub = sp.vstack([A[up], -A[lo]]).tocsr() if (up.any() or lo.any()) else None
    b_ub = np.concatenate([AU[up], -AL[lo]]) if A_ub is not None else None
    A_eq, b_eq = (A[eq], AU[eq]) if eq.any() else (None, None)
    res = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, bounds=np.column_stack([lp.l, lp.u]),
                  method="highs", options={"primal_feasibility_tolerance": 1e-9, "dual_feasibility_tolerance": 1e-9})
    return float(res.fun) + float(lp.obj_constant) if res.status == 0 else None


def defaults_leg(dlp, runs=2):
    """The same resident LP through gp_preprocess with the reference's default
    parameters: folding first (colour refinement; it finds no symmetry in the
    generator LP), then presolve with the structural Startup rules on.  The
    piecewise-linear reformulation branches are not built; a request for them
    is reported in unsupported_stages."""
    import gpupre as G

    out = None
    for k in range(runs):
        G.check(G.load().gp_synchronize())
        t = time.perf_counter()
        s = G.preprocess(dlp, G.default_params())
        G.check(G.load().gp_synchronize())
        wall = time.perf_counter() - t
        f = s.folding
        out = {"wall_ms": round(wall * 1e3, 3), "preprocessing_ms": round(s.preprocessing_seconds * 1e3, 3),
               "folding_ms": round(f.seconds * 1e3, 3), "folding_applied": bool(s.folding_applied),
               "folding_rounds": int(f.rounds), "presolve_core_ms": round(s.elapsed_seconds * 1e3, 3),
               "reduced": {"rows": s.reduced_rows, "cols": s.reduced_cols, "nnz": s.reduced_nnz,
                           "iterations": s.iterations},
               "unsupported_stages": int(s.unsupported_stages), "primal_only_reductions": int(s.primal_only_reductions),
               "runs": k + 1, "note": "second run reported (the first warms the folding kernels)"}
        s.free()
    log(f"reference defaults (folding + structural rules): {out}")
    return out


# The reference's published presolve and preprocessing times on H100 PCIe for
# the two largest QAP20 instances, and their reduced dimensions
# (results/raw/benchmarks/QAP20/03_GPU-Preprocessor/logs/
#  tables_3_4_5_qap20_gpu_preprocessor_{tai60b,lipa70a}.log:16-22,27)
PUBLISHED_QAP = {
    60: dict(instance="tai60b", presolve_ms=202.03673, preprocessing_ms=335.16832, rows=26355720, cols=12963600,
             nnz=65235600, reduced=(432120, 6485400, 26359200)),
    70: dict(instance="lipa70a", presolve_ms=366.32473, preprocessing_ms=619.87047, rows=48711040, cols=24014900,
             nnz=120740900, reduced=(686140, 12012350, 48715800)),
}


def published_qap_leg(seed, runs=3):
    """QAP LPs of the shape of the reference's two largest QAP20 instances
    (generate_qap20.jl's layout at n = 60 and 70, seeded flow / distance
    data), presolved with the reference's defaults: folding, then presolve
    with the structural Startup rules.  Set beside the reference's published
    H100 PCIe times; the reduced dimensions must equal the published ones."""
    import numpy as np
    import gpupre as G

    out = {}
    for n, pub in PUBLISHED_QAP.items():
        lp = G.generate_qap(n, seed=seed + n)
        assert (lp.m, lp.n, lp.nnz) == (pub["rows"], pub["cols"], pub["nnz"])
        with G.DeviceLP(lp) as d:
            core, prep, fold = [], [], []
            for k in range(runs + 1):
                G.check(G.load().gp_synchronize())
                s = G.preprocess(d, G.default_params())
                G.check(G.load().gp_synchronize())
                red = (s.reduced_rows, s.reduced_cols, s.reduced_nnz)
                if k > 0:  # the first run warms the kernels
                    core.append(s.elapsed_seconds * 1e3)
                    prep.append(s.preprocessing_seconds * 1e3)
                    fold.append(s.folding.seconds * 1e3)
                s.free()
        want = pub["reduced"]
        same = red[0] == want[0] and red[1] == want[1] and (want[2] is None or red[2] == want[2])
        c, p = float(np.median(core)), float(np.median(prep))
        out[pub["instance"] + "_shape"] = {
            "n": n, "rows": lp.m, "cols": lp.n, "nnz": lp.nnz,
            "reduced": {"rows": red[0], "cols": red[1], "nnz": red[2]}, "reduced_equals_published": same,
            "presolve_ms_p50": round(c, 3), "preprocessing_ms_p50": round(p, 3),
            "folding_ms_p50": round(float(np.median(fold)), 3),
            "published_presolve_ms": pub["presolve_ms"], "published_preprocessing_ms": pub["preprocessing_ms"],
            "presolve_speedup": round(pub["presolve_ms"] / c, 3),
            "preprocessing_speedup": round(pub["preprocessing_ms"] / p, 3),
            "published_hardware": "NVIDIA H100 PCIe"}
        log(f"QAP n={n} ({pub['instance']} shape): {out[pub['instance'] + '_shape']}")
        del lp
    return out


def lp_solver_leg(seed, medium=False):
    """The FP64 first-order solver (gp_lp_solve) and the north star's gate: the
    presolved LP solved, postsolved, and its objective set against HiGHS on the
    original LP (config 1); with `medium`, the 200k x 300k LP (config 2) against
    the solver's own solve of the original."""
    import numpy as np
    import gpupre as G

    def one(name, tol):
        lp = G.generate_workload(name, seed=seed)
        t = time.perf_counter()
        ref = highs_objective(lp) if name == "small" else None
        highs_s = time.perf_counter() - t
        with G.DeviceLP(lp) as d:
            x, y, z, info = G.solve_lp(d, tolerance=tol, time_limit=120.0)
            s = G.preprocess(d, G.parity_params())
            try:
                xr, yr, zr, rinfo = G.solve_lp(s.s.reduced_lp, tolerance=tol, time_limit=120.0)
                xo, _, _ = G.postprocess(s, xr, yr, zr, d)
                red = (s.reduced_rows, s.reduced_cols)
            finally:
                s.free()
        post = float(np.dot(lp.c, xo)) + float(lp.obj_constant)
        target = ref if ref is not None else info.primal_objective

        def leg(i):
            return {"status": G.LP_STATUS[i.status], "iterations": i.iterations, "restarts": i.restarts,
                    "seconds": round(i.seconds, 3), "objective": i.primal_objective,
                    "us_per_iteration": round(1e6 * i.seconds / max(1, i.iterations), 1),
                    "gbps": round(i.bytes_per_iteration * i.iterations / i.seconds / 1e9, 1) if i.seconds else None}

        out = {"rows": lp.m, "cols": lp.n, "nnz": lp.nnz, "kkt_tolerance": tol, "original": leg(info),
               "reduced": dict(leg(rinfo), rows=red[0], cols=red[1]), "postsolved_objective": post,
               "objective_error": abs(post - target) / max(1.0, abs(target)),
               "error_against": "HiGHS on the original LP" if ref is not None else "this solver on the original LP"}
        if ref is not None:
            out.update(highs_objective=ref, highs_seconds=round(highs_s, 2),
                       original_error=abs(info.primal_objective - ref) / max(1.0, abs(ref)))
        log(f"lp solver {name}: {out}")
        return out

    res = {"config1_small": one("small", 1e-8)}
    if medium:
        res["config2_medium"] = one("medium", 1e-6)
    return res


def mps_ingest_leg(seed):
    """MPS ingest of the 200k x 300k LP (config 2) written as free-format MPS:
    host parse (gp_mps_parse) and device assembly (gp_mps_upload: entries to
    HBM, duplicates summed, A and A^T built), against the host-CSR upload."""
    import gpupre as G

    lp = G.generate_workload("medium", seed=seed)
    text = G.write_mps(lp, "medium")
    t = time.perf_counter()
    model = G.parse_mps(text)
    parse_s = time.perf_counter() - t
    try:
        G.check(G.load().gp_synchronize())
        t = time.perf_counter()
        d = model.upload()
        G.check(G.load().gp_synchronize())
        up_s = time.perf_counter() - t
        entries = model.entries
    finally:
        model.free()
    d.free()
    t = time.perf_counter()
    d2 = G.DeviceLP(lp)
    G.check(G.load().gp_synchronize())
    csr_s = time.perf_counter() - t
    d2.free()
    out = {"rows": lp.m, "cols": lp.n, "entries": int(entries), "text_mb": round(len(text) / 1e6, 1),
           "parse_s": round(parse_s, 3), "parse_mb_per_s": round(len(text) / parse_s / 1e6, 1),
           "device_assembly_ms": round(up_s * 1e3, 2), "host_csr_upload_ms": round(csr_s * 1e3, 2)}
    log(f"mps ingest: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large", choices=["small", "medium", "large", "large10", "large_fp64"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-presolve leg")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-serial", action="store_true",
                    help="batched leg without the one-LP-after-another comparison (profiling runs)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-scheduler leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the per-config GPU vs PSLP table")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-recipe verdict leg")
    ap.add_argument("--no-pslp-full", action="store_true", help="skip PSLP on the whole workload")
    ap.add_argument("--no-lp-solver", action="store_true", help="skip the LP-solver gate leg")
    ap.add_argument("--lp-medium", action="store_true", help="add the 200k x 300k LP to the solver leg (~1 min)")
    ap.add_argument("--no-mps", action="store_true", help="skip the MPS ingest leg")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the QAP legs of the reference's published tai60b / lipa70a shapes")
    ap.add_argument("--no-defaults", action="store_true",
                    help="skip the reference-default pipeline leg (folding + structural rules)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) three times, to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings, and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the presolve engine has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import gpupre as G

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # one elementwise kernel: reads n*8 bytes, writes n*8 bytes
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {n * 8} bytes read and written")
        return

    if G.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    t = time.time()
    lp = G.generate_workload(args.workload, seed=args.seed)
    m0, n0, nnz = lp.m, lp.n, lp.nnz
    log(f"generated {args.workload}: {lp.m} x {lp.n}, nnz={lp.nnz} in {time.time() - t:.1f}s")
    t = time.time()
    dlp = G.DeviceLP(lp)
    log(f"uploaded + device transpose in {time.time() - t:.1f}s")
    params = G.parity_params()

    for w in range(args.warmup):
        s = G.preprocess(dlp, params)
        log(f"warmup {w}: {s.reduced_rows} x {s.reduced_cols} nnz={s.reduced_nnz} core={s.elapsed_seconds*1e3:.1f} ms "
            f"iters={s.iterations}")
        s.free()

    # kernel-family HIP events ride along in the timed steps (null stream, where
    # every engine kernel is launched); they add a few microseconds per family scope
    timed_params = params if args.no_kernel_timing else G.parity_params(kernel_timing=1)
    ks = {name: dict(ms=0.0, launches=0, bytes=0.0) for name in G.KERNEL_FAMILIES}
    G.check(G.load().gp_synchronize())
    core = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        s = G.preprocess(dlp, timed_params)
        core.append(s.elapsed_seconds)
        reduced = (s.reduced_rows, s.reduced_cols, s.reduced_nnz, s.iterations,
                   "infeasible" if s.has_infeasible else "unbounded" if s.has_unbounded else "ok")
        for name, v in s.kernel_stats().items():
            for f in ("ms", "launches", "bytes"):
                ks[name][f] += v[f]
        s.free()
    G.check(G.load().gp_synchronize())
    wall = time.perf_counter() - t0
    ms_per_step = wall / max(1, args.steps) * 1e3
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, core p50 {np.median(core)*1e3:.2f} ms")

    # postsolve of a reduced solution (zeros: the replay cost does not depend on
    # the values) through the public postprocess, original model given
    s = G.preprocess(dlp, params)
    rec_count = s.record().types.size
    zx, zy = np.zeros(s.reduced_cols), np.zeros(s.reduced_rows)
    G.check(G.load().gp_synchronize())
    t = time.perf_counter()
    G.postprocess(s, zx, zy, zx, dlp)
    post_ms = (time.perf_counter() - t) * 1e3
    s.free()
    log(f"postsolve: {post_ms:.1f} ms for {rec_count} tape records")
    defaults = None if args.no_defaults else defaults_leg(dlp)

    roofline = None
    families = {}
    if not args.no_kernel_timing:
        for name, v in ks.items():
            if v["launches"]:
                gbps = round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["bytes"] > 0 and v["ms"] > 0 else None
                families[name] = dict(ms=round(v["ms"], 3), launches=v["launches"], gbps=gbps,
                                      frac=round(gbps / HBM_PEAK_GBPS, 4) if gbps is not None else None)
        # dominant family that brackets one kernel (roofline of that kernel)
        sweep = [n for n in ks if ks[n]["launches"] and ks[n]["bytes"] > 0 and n in KERNEL_OF_FAMILY]
        dom = max(sweep, key=lambda n: ks[n]["ms"])
        v = ks[dom]
        per_launch_bytes = v["bytes"] / v["launches"]
        avg_s = v["ms"] * 1e-3 / v["launches"]
        achieved = per_launch_bytes / avg_s / 1e9
        traffic = measured_traffic(KERNEL_OF_FAMILY[dom], args.workload)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "family": dom,
                    "kernel": KERNEL_OF_FAMILY[dom], "algorithmic_bytes_per_launch": round(per_launch_bytes),
                    "avg_launch_ms": round(avg_s * 1e3, 4), "launches": v["launches"]}
        log(f"kernel families: {json.dumps(families)}")

    cpu = None
    restatement = None
    pslp_full = None
    host = host_info()
    if not args.no_cpu_baseline:
        # PSLP, the reference's bundled CPU presolver, on a bounded sample of the
        # same recipe (the largest fraction on which it completes a presolve)
        frac = {"large": 4, "large_fp64": 4, "large10": 1, "medium": 1, "small": 1}.get(args.workload, 1)
        w = dict(G.WORKLOADS[args.workload])
        m, n = w.pop("m") // frac, w.pop("n") // frac
        sample = lp if frac == 1 else G.generate(m, n, args.seed, **w)
        r = pslp_pinned(sample)
        if r is not None:
            cpu = {"value": round(sample.nnz / r["time_presolve"] / 1e6, 3) if r["time_presolve"] > 0 else None,
                   "unit": "Mnnz/s", "cores": r["cores"], "kind": "reference",
                   "sample": f"PSLP 0.0.11 (the reference's bundled CPU presolver, default settings) on a "
                             f"{sample.m}x{sample.n} LP ({sample.nnz} nnz) from the same recipe ("
                             f"{'1/' + str(frac) + ' of the rows and columns' if frac > 1 else 'the workload itself'}"
                             f"): time_presolve {r['time_presolve']:.3f} s (the figure the reference reports), "
                             f"new_presolver + run_presolver {r['wall_seconds']:.3f} s, outcome {r['outcome']} -> "
                             f"{r['rows']}x{r['cols']} nnz={r['nnz']}; pinned to {r['cores']} CPUs (one thread plus "
                             f"one helper during init) of {host['cpu_model']} ({host['nproc']} CPUs on the host)",
                   "ms": round(r["time_presolve"] * 1e3, 1), "wall_ms": round(r["wall_seconds"] * 1e3, 1),
                   "host": host}
            log(f"cpu baseline (PSLP sample): {cpu}")
        if frac > 1 and not args.no_pslp_full:
            r = pslp_pinned(lp)
            if r is not None:
                pslp_full = {"rows": lp.m, "cols": lp.n, "nnz": lp.nnz, "outcome": r["outcome"],
                             "wall_ms": round(r["wall_seconds"] * 1e3, 1),
                             "time_presolve_ms": round(r["time_presolve"] * 1e3, 1),
                             "note": "PSLP on the whole workload; an 'infeasible' outcome is a tolerance artefact "
                                     "(the generator's planted point satisfies every row and bound, "
                                     "tests/test_generator_feasibility.py)"}
                log(f"PSLP on the full LP: {pslp_full}")
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        # the CPU restatement of this engine (the parity oracle) on a bounded
        # sample of the same recipe
        frac = {"large": 40, "large_fp64": 40, "large10": 4, "medium": 4, "small": 1}.get(args.workload, 1)
        w = dict(G.WORKLOADS[args.workload])
        m, n = w.pop("m") // frac, w.pop("n") // frac
        sample = G.generate(m, n, args.seed, **w)
        t = time.perf_counter()
        O.presolve(sample, params)
        dt = time.perf_counter() - t
        restatement = {"value": round(sample.nnz / dt / 1e6, 3), "unit": "Mnnz/s", "cores": 1, "kind": "port",
                       "sample": f"CPU restatement, full presolve of a {m}x{n} LP ({sample.nnz} nnz) from the same "
                                 f"generator recipe, {dt:.2f} s"}
        log(f"cpu restatement: {restatement}")
        if cpu is None:
            cpu = restatement
            restatement = None

    dlp.free()
    del lp
    batched = None if args.no_batched else batched_leg(args.batch_count, args.seed, not args.no_serial)
    learned = None if args.no_learned else learned_leg(args.seed)
    configs = None if args.no_configs else configs_leg(args.seed)
    fp64 = None if (args.no_fp64 or args.workload != "large") else fp64_verdict_leg(args.seed)
    lp_solver = None if args.no_lp_solver else lp_solver_leg(args.seed, args.lp_medium)
    mps = None if args.no_mps else mps_ingest_leg(args.seed)
    published = None if args.no_published else published_qap_leg(args.seed)

    value = nnz / (ms_per_step * 1e-3) / 1e6
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mnnz/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload), "name": args.workload,
                   "rows": m0, "cols": n0, "nnz": nnz, "seed": args.seed, "folding": False,
                   "structure_specific_rules": False,
                   "reduced": {"rows": reduced[0], "cols": reduced[1], "nnz": reduced[2], "iterations": reduced[3],
                               "status": reduced[4]}},
        "presolve_core_ms_p50": round(float(np.median(core)) * 1e3, 3),
        "postsolve_ms": round(post_ms, 2), "postsolve_records": int(rec_count),
        "kernel_families": families, "roofline": roofline, "cpu_baseline": cpu, "cpu_restatement": restatement,
        "batched": batched, "learned_vs_plain": learned, "configs": configs, "large_fp64_verdict": fp64,
        "pslp_full_lp": pslp_full, "lp_solver": lp_solver, "mps_ingest": mps, "reference_defaults": defaults,
        "published_qap": published,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
