#!/usr/bin/env python
"""MLL + gradient evaluations/s of the exact GP at N=65,536, D=16 on one MI355X.

Workload: ``gpjax_amd.workloads.exact_large`` -- RBF ARD kernel, 65,536 points in
16 dimensions, random-Fourier-feature targets (synthetic, seeded).  One step =
one ``objectives.conjugate_mll`` forward plus its backward (the gradient w.r.t.
the 16 lengthscales, variance and obs_stddev), inputs resident in HBM.

The headline value is the mixed-precision path (float32 storage, BF16-split
products on the matrix cores, float64 diagonal blocks / refinement /
log-determinant).  The plain float64 GPU path is timed beside it; its MLL and
gradient are the accuracy reference at this size (the float64 path itself is
checked against the float64 CPU oracle at smaller N in tests/).  The CPU
baseline runs the float64 oracle on the host cores on bounded samples (same
recipe at N = 2,048, 4,096, 8,192), fits t = c N^3 and extrapolates to N=65,536.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("MLL+grad evals/s at N=65,536: mixed-precision vs FP64 GPU, "
          "at |ΔMLL|/|MLL|≤1e-6")
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def lauum_flops(n: int, tile: int = 128) -> float:
    """Algorithmic flops of the fused Sigma^-1 = X^T X pass (lower tiles, k >= row tile)."""
    nt = (n + tile - 1) // tile
    return 2.0 * tile ** 3 * nt * (nt + 1) * (nt + 2) / 6.0


def profiled_traffic(n: int):
    """HBM bytes per launch of the adjoint kernel at N = 65,536, from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary (scaled by the copy calibration)."""
    if n != 65536:
        return None
    path = os.path.join(ROOT, "profiles", "adjoint_bf16x6_n65536_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        return None
    return None


def run_steps(post, D, precision, steps, params, report=False):
    """`steps` MLL + gradient evaluations; with report=True each step syncs and
    prints its wall time (used for the untimed warm-up steps only)."""
    import torch

    import gpjax_amd as gpx
    out = None
    for i in range(steps):
        t0 = time.perf_counter()
        for p in params:
            p.value.grad = None
        mll, res = gpx.objectives.conjugate_mll_with_info(post, D, precision=precision)
        mll.backward()
        out = (mll, res)
        if report:
            torch.cuda.synchronize()
            print(f"# {precision} warm-up step {i}: {time.perf_counter() - t0:.2f}s "
                  f"mll={float(mll.detach()):.6f}", file=sys.stderr, flush=True)
    return out


def grads(params):
    import torch
    return torch.cat([p.value.grad.reshape(-1) for p in params]).double().cpu()


def cpu_baseline(n: int, workload: str = "exact_large") -> dict:
    threads = os.cpu_count() or 1
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
               MKL_NUM_THREADS=str(threads))
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--n", str(n),
           "--threads", str(threads), "--workload", workload]
    if workload == "exact_large":
        cmd += ["--ns", "2048,4096,8192"]
    proc = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if proc.returncode != 0:
        return {"value": None, "error": proc.stderr[-500:]}
    return json.loads(proc.stdout.strip().splitlines()[-1])


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def batched_mixed_line(gpx, w, dev, args, t64, mll64) -> dict:
    """The many-small-GPs batch on the mixed path (float32 storage, bf16 splits,
    float64 panels, each problem's own probe check), timed with the check's host
    read off; compared with the float64 MLLs of the same parameters."""
    mm = gpx.batched.BatchedConjugateMLL(w["Xs"], w["ys"], kernel="rbf",
                                         lengthscale=w["lengthscale"], variance=w["variance"],
                                         obs_stddev=w["obs_stddev"], jitter=w["jitter"],
                                         device=dev, precision="mixed")
    params = [mm.lengthscale, mm.variance, mm.obs_stddev]

    def step():
        for p in params:
            p.value.grad = None
        mll = mm.mll()
        mll.sum().backward()
        return mll

    old_check = gpx.config.mixed_check
    gpx.config.mixed_check = "off"
    try:
        t_mx, mll_mx = timed(step, args.steps, args.warmup)
    finally:
        gpx.config.mixed_check = old_check
    kept = mm.last_result[:, 7] == 0.0
    mll_mx, mll64 = mll_mx.detach(), mll64.detach()
    rel = ((mll_mx - mll64).abs() / mll64.abs())[kept]
    return {"ms_per_step": round(t_mx * 1e3, 2), "speedup_vs_fp64": round(t64 / t_mx, 3),
            "flagged_problems": int((~kept).sum()),
            "max_rel_mll_diff_unflagged_vs_fp64": float(rel.max()) if rel.numel() else None,
            "note": "with any problem flagged, mixed_check='fallback' recomputes the batch "
                    "in float64"}


def batched_fit_line(gpx, w, dev, args) -> dict:
    """Training: Adam steps on the summed negative MLL of the batch, one step
    (objective, backward, fused device Adam) captured in a HIP graph and replayed."""
    import torch
    fm = gpx.batched.BatchedConjugateMLL(w["Xs"], w["ys"], kernel="rbf",
                                         lengthscale=w["lengthscale"], variance=w["variance"],
                                         obs_stddev=w["obs_stddev"], jitter=w["jitter"],
                                         device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = gpx.fit(model=fm, objective=lambda mdl, data: -mdl.mll().sum(), train_data=None,
                      num_iters=args.fit_steps, learning_rate=0.01, capture_graph=True,
                      warmup=3)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    hist = hist.cpu()
    return {"adam_steps": args.fit_steps, "graph_captured": True, "seconds": round(t_fit, 3),
            "ms_per_step": round(t_fit / args.fit_steps * 1e3, 2),
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "note": "includes 3 eager warm-up steps and the capture"}


def other_workload(args) -> dict:
    """The other configurations, one JSON line each (not the headline):
    matern  -- exact MLL + gradient, N = 16,384, D = 8, Matern-5/2 ARD, mixed and FP64;
    loocv   -- conjugate_loocv + gradient on the same data, mixed and FP64;
    batched -- 4,096 independent exact GPs (N_i in 32..1024, D = 4, RBF ARD), MLL +
               gradient of all of them per step;
    sparse  -- collapsed ELBO + gradient (theta and the 2,048 x 8 inducing inputs),
               N = 1e6, D = 8, Matern-3/2 ARD, mixed and FP64;
    sparse_predict -- CollapsedVariationalGaussian.predict, diagonal variance, on the
               sparse workload's data and inducing inputs at T = 1,000,000, FP64 (the
               ELBO's statistics recomputed in every step);
    predict -- ConjugatePosterior.predict with the diagonal variance at large T:
               the matern workload's N = 16,384 training points, T = 1,000,000 test
               points ~ N(0, I_8), mixed and FP64 (factorisation included in every
               step)."""
    import torch

    import gpjax_amd as gpx
    from gpjax_amd import workloads
    dev = torch.device("cuda:0")
    line = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    if args.workload == "batched":
        w = workloads.batched()
        m = gpx.batched.BatchedConjugateMLL(w["Xs"], w["ys"], kernel="rbf",
                                            lengthscale=w["lengthscale"], variance=w["variance"],
                                            obs_stddev=w["obs_stddev"], jitter=w["jitter"],
                                            device=dev)
        params = [m.lengthscale, m.variance, m.obs_stddev]

        def step():
            for p in params:
                p.value.grad = None
            mll = m.mll()
            mll.sum().backward()
            return mll

        t, mll = timed(step, args.steps, args.warmup)
        sizes = [int(s) for s in w["sizes"]]
        line["autotune"] = [{"padded_size": r["padded_size"], "members": r["members"],
                             "stored_sigma": r["stored_sigma"],
                             "transposed_pass": r["transposed_pass"],
                             "ms": {k: round(v, 3) for k, v in r["ms"].items()}}
                            for r in m.tune_report]
        cpu_n = 32
        line.update(metric="batched MLL+grad evals/s (4096 exact GPs, N_i in 32..1024, D=4)",
                    value=round(1.0 / t, 4), unit="evals/s", ms_per_step=round(t * 1e3, 2),
                    dtype="f64", config={"workload": "batched: 4096 RBF-ARD GPs "
                                         "(gpjax_amd.workloads.batched)",
                                         "problems": len(sizes), "sum_n": sum(sizes),
                                         "sum_n3": float(sum(s ** 3 for s in sizes))},
                    all_finite=bool(torch.isfinite(mll).all()))
        if not args.eval_only:
            line["mixed"] = batched_mixed_line(gpx, w, dev, args, t, mll)
            line["fit"] = batched_fit_line(gpx, w, dev, args)
    elif args.workload == "sparse":
        w = workloads.sparse()
        k = gpx.kernels.Matern32(lengthscale=w["lengthscale"], variance=w["variance"])
        post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k) * \
            gpx.likelihoods.Gaussian(num_datapoints=len(w["X"]), obs_stddev=w["obs_stddev"])
        q = gpx.variational_families.CollapsedVariationalGaussian(posterior=post,
                                                                 inducing_inputs=w["Z"])
        D = gpx.Dataset(X=torch.tensor(w["X"], device=dev),
                        y=torch.tensor(w["y"][:, None], device=dev))
        params = [k.lengthscale, k.variance, post.likelihood.obs_stddev, q.inducing_inputs]

        def make(prec):
            def step():
                for p in params:
                    p.value.grad = None
                e, res = gpx.objectives.collapsed_elbo_with_info(q, D, precision=prec)
                e.backward()
                return float(e.detach()), grads(params), res
            return step

        t_m, (e_m, g_m, res_m) = timed(make("mixed"), args.steps, args.warmup)
        t_d, (e_d, g_d, _) = timed(make("fp64"), args.steps, args.warmup)
        nth = g_d.numel() - w["Z"].size  # theta entries first, then the 2,048 x 8 Z gradient
        gt_m, gt_d, gz_m, gz_d = g_m[:nth], g_d[:nth], g_m[nth:], g_d[nth:]
        res_m = res_m.cpu()
        flag = float(res_m[gpx._lib.RES_FLAG])
        line.update(metric="collapsed ELBO+grad evals/s at N=1e6, M=2048, mixed",
                    value=round(1.0 / t_m, 4), unit="evals/s", ms_per_step=round(t_m * 1e3, 2),
                    dtype="f32 Kzx; P as bf16x6-split, gradient pass as fp16x3-split "
                          "products with f64 sums; f64 m x m algebra",
                    config={"workload": "sparse: N=1e6, D=8, M=2048, Matern-3/2 ARD "
                                        "(gpjax_amd.workloads.sparse)", "n": len(w["X"]),
                            "m": 2048, "d": 8}, elbo=e_m,
                    fp64_gpu={"evals_per_s": round(1.0 / t_d, 4),
                              "ms_per_step": round(t_d * 1e3, 2), "elbo": e_d},
                    speedup_vs_fp64=round(t_d / t_m, 3),
                    accuracy={"rel_elbo_vs_fp64": abs(e_m - e_d) / abs(e_d),
                              "rel_grad_theta_vs_fp64":
                                  float(torch.linalg.norm(gt_m - gt_d) / torch.linalg.norm(gt_d)),
                              "rel_grad_z_vs_fp64":
                                  float(torch.linalg.norm(gz_m - gz_d) / torch.linalg.norm(gz_d)),
                              "max_grad_z_err_of_max_entry":
                                  float((gz_m - gz_d).abs().max() / gz_d.abs().max())},
                    mixed_check={"flag": flag,
                                 "elbo_err_est": float(res_m[gpx._lib.RES_MLL_ERR]),
                                 "meaning": "flag 0: mixed result kept; 1: flagged; 2: "
                                            "recomputed in float64 (then the timed steps are "
                                            "not the mixed path)"})
        cpu_n = 16384
    elif args.workload == "predict":
        import numpy as np
        w = workloads.exact_matern()
        n, t = len(w["X"]), 1_000_000
        k = gpx.kernels.Matern52(lengthscale=w["lengthscale"], variance=w["variance"])
        post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k,
                             jitter=w["jitter"]) * \
            gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=w["obs_stddev"])
        D = gpx.Dataset(X=torch.tensor(w["X"], device=dev),
                        y=torch.tensor(w["y"][:, None], device=dev))
        T = torch.tensor(np.random.default_rng(7).standard_normal((t, 8)), device=dev)

        def make(prec):
            def step():
                return post.predict(T, D, full_covariance=False, precision=prec)
            return step

        tm, dist = timed(make("mixed"), args.steps, args.warmup)
        td, dist64 = timed(make("fp64"), args.steps, args.warmup)
        rel = lambda a, b: float((a - b).abs().max() / b.abs().max())  # noqa: E731
        line.update(metric="test points/s, diagonal predictive, N=16,384, T=1e6, mixed",
                    value=round(t / tm, 1), unit="test points/s", ms_per_step=round(tm * 1e3, 2),
                    dtype="f32 storage, bf16x6-split products, f64 panels and refinement",
                    config={"workload": "predict: matern training data, T=1e6 test points "
                                        "(gpjax_amd.workloads.exact_matern)", "n": n, "t": t,
                            "d": 8},
                    fp64_gpu={"test_points_per_s": round(t / td, 1),
                              "ms_per_step": round(td * 1e3, 2)},
                    speedup_vs_fp64=round(td / tm, 3),
                    accuracy={"mean_max_err_of_max_vs_fp64": rel(dist.mean, dist64.mean),
                              "variance_max_err_of_max_vs_fp64":
                                  rel(dist.variance, dist64.variance)},
                    all_finite=bool(torch.isfinite(dist.mean).all()
                                    and torch.isfinite(dist.variance).all()))
        cpu_n = None
    elif args.workload == "sparse_predict":
        import numpy as np
        w = workloads.sparse()
        t = 1_000_000
        k = gpx.kernels.Matern32(lengthscale=w["lengthscale"], variance=w["variance"])
        post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k) * \
            gpx.likelihoods.Gaussian(num_datapoints=len(w["X"]), obs_stddev=w["obs_stddev"])
        q = gpx.variational_families.CollapsedVariationalGaussian(posterior=post,
                                                                 inducing_inputs=w["Z"])
        D = gpx.Dataset(X=torch.tensor(w["X"], device=dev),
                        y=torch.tensor(w["y"][:, None], device=dev))
        T = torch.tensor(np.random.default_rng(11).uniform(0.0, 1.0, (t, 8)), device=dev)

        def step():
            return q.predict(T, D, full_covariance=False)

        tm, dist = timed(step, args.steps, args.warmup)
        line.update(metric="test points/s, collapsed diagonal predictive, N=1e6, M=2048, "
                           "T=1e6 (FP64)",
                    value=round(t / tm, 1), unit="test points/s", ms_per_step=round(tm * 1e3, 2),
                    dtype="f64",
                    config={"workload": "sparse_predict: C4 training data and inducing inputs, "
                                        "T=1e6 test points ~ U[0,1]^8 "
                                        "(gpjax_amd.workloads.sparse)", "n": len(w["X"]),
                            "m": 2048, "t": t, "d": 8},
                    all_finite=bool(torch.isfinite(dist.mean).all()
                                    and torch.isfinite(dist.variance).all()))
        cpu_n = None
    elif args.workload == "loocv":
        w = workloads.exact_matern()
        n = len(w["X"])
        k = gpx.kernels.Matern52(lengthscale=w["lengthscale"], variance=w["variance"])
        post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k,
                             jitter=w["jitter"]) * \
            gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=w["obs_stddev"])
        D = gpx.Dataset(X=torch.tensor(w["X"], device=dev),
                        y=torch.tensor(w["y"][:, None], device=dev))
        params = [k.lengthscale, k.variance, post.likelihood.obs_stddev]

        def make(prec):
            def step():
                for p in params:
                    p.value.grad = None
                obj = gpx.objectives.conjugate_loocv(post, D, precision=prec)
                obj.backward()
                return float(obj.detach()), grads(params)
            return step

        t_m, (v_m, g_m) = timed(make("mixed"), args.steps, args.warmup)
        t_d, (v_d, g_d) = timed(make("fp64"), args.steps, args.warmup)
        line.update(metric="LOOCV+grad evals/s at N=16,384 (Matern-5/2 ARD, D=8), mixed",
                    value=round(1.0 / t_m, 4), unit="evals/s", ms_per_step=round(t_m * 1e3, 2),
                    dtype="bf16x6-split/f32 with f64 panels",
                    config={"workload": "conjugate_loocv + grad, N=16384 D=8 Matern-5/2 ARD "
                                        "(gpjax_amd.workloads.exact_matern)", "n": n, "d": 8},
                    fp64_gpu={"evals_per_s": round(1.0 / t_d, 4),
                              "ms_per_step": round(t_d * 1e3, 2)},
                    speedup_vs_fp64=round(t_d / t_m, 3),
                    accuracy={"rel_loocv_vs_fp64": abs(v_m - v_d) / abs(v_d),
                              "rel_grad_vs_fp64":
                                  float(torch.linalg.norm(g_m - g_d) / torch.linalg.norm(g_d))})
        cpu_n = 4096
    else:  # matern
        w = workloads.exact_matern()
        n = len(w["X"])
        k = gpx.kernels.Matern52(lengthscale=w["lengthscale"], variance=w["variance"])
        post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k,
                             jitter=w["jitter"]) * \
            gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=w["obs_stddev"])
        D = gpx.Dataset(X=torch.tensor(w["X"], device=dev),
                        y=torch.tensor(w["y"][:, None], device=dev))
        params = [k.lengthscale, k.variance, post.likelihood.obs_stddev]

        def make(prec):
            def step():
                for p in params:
                    p.value.grad = None
                mll = gpx.objectives.conjugate_mll(post, D, precision=prec)
                mll.backward()
                return float(mll.detach()), grads(params)
            return step

        t_m, (mll_m, g_m) = timed(make("mixed"), args.steps, args.warmup)
        t_d, (mll_d, g_d) = timed(make("fp64"), args.steps, args.warmup)
        rel_g = float(torch.linalg.norm(g_m - g_d) / torch.linalg.norm(g_d))
        line.update(metric="MLL+grad evals/s at N=16,384 (Matern-5/2 ARD, D=8), mixed",
                    value=round(1.0 / t_m, 4), unit="evals/s", ms_per_step=round(t_m * 1e3, 2),
                    dtype="bf16x6-split/f32 with f64 panels",
                    config={"workload": "exact GP MLL+grad, N=16384 D=8 Matern-5/2 ARD "
                                        "(gpjax_amd.workloads.exact_matern)", "n": n, "d": 8},
                    fp64_gpu={"evals_per_s": round(1.0 / t_d, 4),
                              "ms_per_step": round(t_d * 1e3, 2)},
                    speedup_vs_fp64=round(t_d / t_m, 3),
                    accuracy={"rel_mll_vs_fp64": abs(mll_m - mll_d) / abs(mll_d),
                              "rel_grad_vs_fp64": rel_g})
        cpu_n = 4096
    line["cpu_baseline"] = None if (args.no_cpu or cpu_n is None) else \
        cpu_baseline(cpu_n, args.workload)
    return line


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--fp64-warmup", type=int, default=1)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--fit-steps", type=int, default=100,
                    help="batched workload: graph-captured Adam steps timed after the evaluations")
    ap.add_argument("--eval-only", action="store_true",
                    help="batched workload: time the float64 evaluations only (no mixed-path "
                         "line, no Adam steps), e.g. for a kernel trace")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-arith", default="6,3,6", metavar="P,T,L",
                    help="bf16 products per element for the Cholesky updates, the "
                         "triangular inverse and the fused Sigma^-1 pass (0 = float64 "
                         "products, 3, 4 = fp16 hi/lo split, or 6 each)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes 5 times "
                         "(reference for rocprof FETCH_SIZE / WRITE_SIZE scaling) and exit")
    ap.add_argument("--workload", default="exact_large",
                    choices=["exact_large", "matern", "loocv", "batched", "sparse",
                             "predict", "sparse_predict"],
                    help="exact_large (default) is the headline; the others print their own line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): gpjax_amd has no "
              "multi-GPU path", file=sys.stderr)
        return 2
    if args.workload != "exact_large":
        print(json.dumps(other_workload(args)), flush=True)
        return 0
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "launches": 5}))
        return 0

    import numpy as np
    import torch

    import gpjax_amd as gpx
    from gpjax_amd import _lib, workloads

    dev = torch.device("cuda:0")
    arith = [int(v) for v in args.mixed_arith.split(",")]
    mixed = "mixed:%d,%d,%d" % tuple(arith)
    _lib.mixed_arith(*arith)  # validates the stages
    # The a-posteriori check flags every result whose Cholesky or fused-pass
    # stage is outside its calibration (fp16 or three-product splits); timing
    # those stages means timing them unchecked (the flag is still reported).
    calibrated = arith[0] in (0, 6) and arith[2] in (0, 6)
    if not calibrated:
        gpx.config.mixed_check = "off"
    t0 = time.time()
    w = workloads.exact_large(n=args.n)
    print(f"# workload generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    X = torch.tensor(w["X"], device=dev)
    y = torch.tensor(w["y"][:, None], device=dev)
    D = gpx.Dataset(X=X, y=y)
    kern = gpx.kernels.RBF(lengthscale=w["lengthscale"], variance=w["variance"])
    lik = gpx.likelihoods.Gaussian(num_datapoints=args.n, obs_stddev=w["obs_stddev"])
    post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=kern,
                         jitter=w["jitter"]) * lik
    params = [kern.lengthscale, kern.variance, lik.obs_stddev]

    # ---- mixed precision: the headline ----
    run_steps(post, D, mixed, args.warmup, params, report=True)
    torch.cuda.synchronize()
    _lib.profile_collect()
    _lib.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mll_m, res_m = run_steps(post, D, mixed, args.steps, params)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    _lib.profile_enable(False)
    phases_m = _lib.profile_collect()
    g_m = grads(params)
    mll_m = float(mll_m.detach())
    resid_m = float(res_m[_lib.RES_RESID])
    probe_m = float(res_m[_lib.RES_PROBE])
    flag_m = float(res_m[_lib.RES_FLAG])
    print(f"# mixed: {t_mixed / args.steps:.3f}s/eval mll={mll_m:.6f} resid={resid_m:.2e} "
          f"phases={phases_m}", file=sys.stderr, flush=True)

    # ---- plain float64 GPU path ----
    fp64 = None
    if not args.no_fp64:
        run_steps(post, D, "fp64", args.fp64_warmup, params, report=True)
        torch.cuda.synchronize()
        _lib.profile_collect()
        _lib.profile_enable(True)
        t0 = time.perf_counter()
        mll_d, res_d = run_steps(post, D, "fp64", args.fp64_steps, params)
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        _lib.profile_enable(False)
        phases_d = _lib.profile_collect()
        g_d = grads(params)
        mll_d = float(mll_d.detach())
        fp64 = dict(t=t_fp64 / args.fp64_steps, mll=mll_d, g=g_d, phases=phases_d,
                    resid=float(res_d[_lib.RES_RESID]))
        print(f"# fp64: {fp64['t']:.3f}s/eval mll={mll_d:.6f} phases={phases_d}",
              file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (fused Sigma^-1 + adjoint, one launch per eval) ----
    # Algorithmic work: the LAUUM X^T X over the lower tiles (n^3/3).  The bf16
    # split issues `arith[2]` products per algorithmic product; that issue rate is
    # reported beside the roofline fraction, not as it.
    adj_ms, adj_cnt = phases_m["adjoint"]
    adj_s = adj_ms / max(adj_cnt, 1) / 1e3
    flops = lauum_flops(args.n)
    achieved = flops / adj_s / 1e12
    issued = max(arith[2], 1) * flops
    traffic = profiled_traffic(args.n) if tuple(arith) == (6, 3, 6) else None
    roofline = {"kernel": "adjoint (fused L^-T L^-1 tile GEMM + gradient epilogue, "
                          f"bf16x{arith[2]})",
                "bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 4),
                "traffic": traffic, "traffic_unit": "bytes per launch (HBM)",
                "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, scaled by "
                                   "a 1 GiB copy's counters; read from the committed "
                                   "profiles/adjoint_bf16x6_n65536_traffic.md, not measured "
                                   "in this run") if traffic else None,
                "algorithmic_flops_per_launch": flops, "avg_launch_ms": round(adj_s * 1e3, 3),
                "issued_bf16_flops_per_launch": issued,
                "mfma_issue_frac": round(issued / adj_s / 1e12 / BF16_PEAK_TFLOPS, 4),
                "eval_frac_of_bf16_peak": None}

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)

    value = args.steps / t_mixed
    # whole evaluation: n^3 algorithmic flops (POTRF + TRTRI + LAUUM, n^3/3 each)
    roofline["eval_frac_of_bf16_peak"] = round(float(args.n) ** 3 * value / 1e12
                                               / BF16_PEAK_TFLOPS, 4)
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mixed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x6-split/f32 with f64 panels", "data": "synthetic",
        "config": {"workload": "exact GP MLL+grad, N=65536 D=16 RBF-ARD, RFF targets "
                               "(gpjax_amd.workloads.exact_large)",
                   "n": args.n, "d": 16, "params": 18, "precision": "mixed",
                   "bf16_products_potrf_trtri_lauum": arith},
        "phases_ms": {k: round(v[0] / max(v[1], 1), 2) for k, v in phases_m.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        rel_mll = abs(mll_m - fp64["mll"]) / abs(fp64["mll"])
        rel_g = float(torch.linalg.norm(g_m - fp64["g"]) / torch.linalg.norm(fp64["g"]))
        line["fp64_gpu"] = {"evals_per_s": round(1.0 / fp64["t"], 4),
                            "ms_per_step": round(fp64["t"] * 1e3, 1),
                            "phases_ms": {k: round(v[0] / max(v[1], 1), 2)
                                          for k, v in fp64["phases"].items()},
                            "tflops_fp64_equiv": round(
                                (args.n ** 3) / fp64["t"] / 1e12, 2)}
        line["speedup_vs_fp64"] = round(fp64["t"] / (t_mixed / args.steps), 3)
        line["accuracy"] = {"rel_mll_vs_fp64": rel_mll, "rel_grad_vs_fp64": rel_g,
                            "mll": mll_m, "mll_fp64": fp64["mll"],
                            "refine_resid_mixed": resid_m,
                            "gate_mll_1e-6": rel_mll <= 1e-6, "gate_grad_1e-4": rel_g <= 1e-4}
    # the mixed path's own a-posteriori check (Rademacher probes of I - Sigma X^T X)
    line["mixed_check"] = {"action": gpx.config.mixed_check,
                           "probe_frob_norm": probe_m * math.sqrt(args.n),
                           "sinv_frob_est": float(res_m[_lib.RES_SINV]),
                           "mll_err_est": float(res_m[_lib.RES_MLL_ERR]),
                           "flag": flag_m, "flag_threshold": 0.75,
                           "meaning": "flag 0: mixed result kept; 1: flagged; 2: recomputed in "
                                      "float64 (then the timed steps are not the mixed path)"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
