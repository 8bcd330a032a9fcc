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


def _median_time(fn, runs):
    fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def other_workload(a) -> dict:
    """Bounded samples of the other configurations, extrapolated to full size:
    matern and loocv (N = 16,384, D = 8) by N^3; batched (4,096 GPs) by the sum of
    N_i^3 over the problems timed against the whole batch; sparse (N = 1e6,
    M = 2,048) linearly in N (its cost is M^2 N + O(M^3))."""
    import numpy as np

    from gpjax_amd import workloads
    from oracle import gp_oracle as O

    if a.workload == "matern":
        w = workloads.exact_matern(n=a.n)
        t = _median_time(lambda: O.conjugate_mll_grad(
            w["kind"], w["X"], w["y"], w["lengthscale"], w["variance"], w["obs_stddev"],
            w["jitter"]), a.runs)
        scale = (a.n / 16384) ** 3
        sample = f"float64 oracle MLL + gradient on exact_matern with N={a.n}; rate scaled to N=16384 by N^3"
    elif a.workload == "loocv":
        w = workloads.exact_matern(n=a.n)
        t = _median_time(lambda: O.conjugate_loocv_grad(
            w["kind"], w["X"], w["y"], w["lengthscale"], w["variance"], w["obs_stddev"],
            w["jitter"]), a.runs)
        scale = (a.n / 16384) ** 3
        sample = (f"float64 oracle LOOCV + gradient on exact_matern with N={a.n}; rate scaled "
                  "to N=16384 by N^3")
    elif a.workload == "batched":
        w = workloads.batched()
        k = a.n  # number of problems timed

        def run():
            for p in range(k):
                O.conjugate_mll_grad("rbf", w["Xs"][p], w["ys"][p], w["lengthscale"][p],
                                     w["variance"][p], w["obs_stddev"][p], w["jitter"])
        t = _median_time(run, a.runs)
        sizes = np.asarray(w["sizes"], dtype=np.float64)
        scale = float((sizes[:k] ** 3).sum() / (sizes ** 3).sum())
        sample = (f"float64 oracle MLL + gradient of the first {k} of the 4096 problems; rate "
                  "scaled to the whole batch by sum(N_i^3)")
    else:  # sparse
        w = workloads.sparse(n=a.n, m=2048)
        t = _median_time(lambda: O.collapsed_elbo_grad(
            w["kind"], w["X"], w["y"], w["Z"], w["lengthscale"], w["variance"],
            w["obs_stddev"], w["jitter"]), a.runs)
        scale = a.n / 1_000_000
        sample = (f"float64 oracle collapsed ELBO + gradient (theta and Z) with N={a.n}, M=2048; "
                  "rate scaled to N=1e6 linearly")
    return {"value": 1.0 / t * scale, "unit": "evals/s", "cores": a.threads, "kind": "port",
            "sample": f"{sample}; median {t:.3f}s of {a.runs} runs after 1 warm-up, "
                      f"{a.threads} threads", "sample_seconds_per_eval": t}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=6144)
    ap.add_argument("--ns", default="2048,4096,8192",
                    help="exact_large: sample sizes whose times are fitted by c * N^3")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--target-n", type=int, default=65536)
    ap.add_argument("--workload", default="exact_large",
                    choices=["exact_large", "matern", "loocv", "batched", "sparse"])
    a = ap.parse_args()
    if a.workload != "exact_large":
        print(json.dumps(other_workload(a)))
        return

    import numpy as np  # noqa: F401  (BLAS threads fixed by the caller's environment)

    from gpjax_amd import workloads
    from oracle import gp_oracle as O

    ns = [int(v) for v in a.ns.split(",")]
    times = []
    for i, n in enumerate(ns):
        w = workloads.exact_large(n=n)
        args = (w["kind"], w["X"], w["y"], w["lengthscale"], w["variance"], w["obs_stddev"],
                w["jitter"])
        if i == 0:
            O.conjugate_mll_grad(*args)  # warm-up (BLAS threads, page faults)
        t0 = time.perf_counter()
        O.conjugate_mll_grad(*args)
        times.append(time.perf_counter() - t0)
    # t = c N^3 with c the mean of t / N^3: least squares on the relative residuals
    c = sum(t / n ** 3 for n, t in zip(ns, times)) / len(ns)
    t_target = c * a.target_n ** 3
    print(json.dumps({
        "value": 1.0 / t_target, "unit": "evals/s", "cores": a.threads, "kind": "port",
        "sample": (f"float64 oracle (LU slogdet/solve + explicit inverse for the gradient) on "
                   f"exact_large, D=16, at N={ns}: one timed run each after one warm-up, "
                   f"{a.threads} threads; times {[round(t, 3) for t in times]} s fitted by "
                   f"t = c N^3 (c = {c:.4e} s), rate extrapolated to N={a.target_n}"),
        "sample_seconds": [round(t, 4) for t in times], "fit_c_seconds_per_n3": c,
        "fit_max_rel_residual": max(abs(t - c * n ** 3) / (c * n ** 3) for n, t in zip(ns, times)),
    }))


if __name__ == "__main__":
    main()
