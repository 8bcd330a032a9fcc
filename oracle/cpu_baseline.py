"""CPU baseline: the float64 oracle's MLL + gradient on the host cores.

TEST / BENCHMARK INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg as a
subprocess, so the BLAS thread count is fixed before NumPy loads).  It times
``gp_oracle.conjugate_mll_grad`` -- the LU-based slogdet/solve path GPJax takes
on CPU plus the explicit inverse that ``jax.grad`` forms -- on the same
workload recipe at a bounded N (``workloads.exact_large(n)``), takes the median
of 3 runs after one warm-up, and extrapolates the rate to N = 65,536 by N^3.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=6144)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--target-n", type=int, default=65536)
    a = ap.parse_args()

    import numpy as np  # noqa: F401  (BLAS threads fixed by the caller's environment)

    from gpjax_amd import workloads
    from oracle import gp_oracle as O

    w = workloads.exact_large(n=a.n)
    args = (w["kind"], w["X"], w["y"], w["lengthscale"], w["variance"], w["obs_stddev"],
            w["jitter"])
    O.conjugate_mll_grad(*args)
    times = []
    for _ in range(a.runs):
        t0 = time.perf_counter()
        O.conjugate_mll_grad(*args)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    scale = (a.n / a.target_n) ** 3
    print(json.dumps({
        "value": 1.0 / t * scale, "unit": "evals/s", "cores": a.threads, "kind": "port",
        "sample": (f"float64 oracle (LU slogdet/solve + explicit inverse for the gradient) on "
                   f"exact_large with N={a.n}, D=16: median {t:.3f}s of {a.runs} runs after 1 "
                   f"warm-up, {a.threads} threads; rate extrapolated to N={a.target_n} by N^3"),
        "sample_seconds_per_eval": t,
    }))


if __name__ == "__main__":
    main()
