#!/usr/bin/env python3
"""A fixed number of iterations of the device LP solver on one of the bench
workloads, for kernel profiles of the solver:

  lp_profile.py <workload> <iterations>

(the tolerance is set below any reachable value, so the run does exactly the
iterations asked for, rounded up to the 64-iteration check window)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import gpupre as G  # noqa: E402


def main():
    name = sys.argv[1] if len(sys.argv) > 1 else "medium"
    iters = int(sys.argv[2]) if len(sys.argv) > 2 else 3008
    lp = G.generate_workload(name, seed=0)
    with G.DeviceLP(lp) as d:
        x, y, z, info = G.solve_lp(d, max_iterations=iters, tolerance=1e-30)
    print(f"[lp] {name}: {info.iterations} iterations in {info.seconds:.3f} s, "
          f"{1e6 * info.seconds / info.iterations:.1f} us per iteration, "
          f"{info.bytes_per_iteration / 1e6:.1f} MB of algorithmic traffic per iteration", flush=True)


if __name__ == "__main__":
    main()
