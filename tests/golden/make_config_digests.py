"""Writes tests/golden/config_digests.json: the CPU restatement's results on
the benchmark configurations too large to rerun inside a test, reduced to
digests (tests/digests.py: counts, the schedule trace, status flags, the
objective constant's bits and SHA-256 of every map, tape and reduced-LP
array), for tests/test_gpu_configs.py to compare the engine against.

* workloads: the full 5M x 8M presolve of the integer recipe `large` and the
  FP64 recipe `large_fp64` (seed 0, parity parameters, schedule trace on);
* batch: 64 instances of the batched workload (rows log-uniform in
  [1e3, 1e5], FP64 data, seed 0), presolved as their host direct sum.

Run from the repository root:
  python tests/golden/make_config_digests.py [large|large_fp64|batch ...]
(no argument: all three; an existing file keeps the entries not rerun).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "gpu-preprocessor-ijoc_amd"), os.path.join(REPO, "oracle"),
                os.path.join(REPO, "tests")]

import gpupre as G  # noqa: E402
import oracle as O  # noqa: E402
from digests import digest_oracle  # noqa: E402

PATH = os.path.join(HERE, "config_digests.json")
BATCH_COUNT, BATCH_SEED = 64, 0


def main(which):
    out = {"about": "CPU restatement digests of the benchmark configurations (parity parameters, schedule trace "
                    "on); written by tests/golden/make_config_digests.py", "workloads": {}}
    if os.path.exists(PATH):
        with open(PATH) as f:
            out = json.load(f)
    for name in which:
        t = time.time()
        if name == "batch":
            lps = G.batch_workload(BATCH_COUNT, seed=BATCH_SEED, integral=0)
            lp = G.direct_sum(lps)
            res = O.presolve(lp, G.parity_params(collect_trace=1))
            d = digest_oracle(res)
            d.update(count=BATCH_COUNT, seed=BATCH_SEED, rows=lp.m, cols=lp.n, nnz=lp.nnz)
            out["batch"] = d
        else:
            lp = G.generate_workload(name, seed=0)
            res = O.presolve(lp, G.parity_params(collect_trace=1))
            d = digest_oracle(res)
            d.update(rows=lp.m, cols=lp.n, nnz=lp.nnz)
            out["workloads"][name] = d
        d["seconds"] = round(time.time() - t, 1)
        print(name, {k: v for k, v in d.items() if k not in ("sha256", "trace")}, flush=True)
        res.free()
        with open(PATH, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or ["large", "large_fp64", "batch"])
