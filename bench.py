"""Benchmark of the headline workload: fused forward+adjoint KL gradient.

Workload (config C3): 8-D triangular map, total order 3 per component
(494 coefficients), ProbabilistHermite / SoftPlus / adaptive Simpson
(MapOptions defaults), N = 500,000 FP64 points from the banana recursion
x1 ~ N(0,1), x_k = 0.5 x_{k-1}^2 + tanh(x_{k-2}) + 0.3 N(0,1) (seed 0).
Coefficients ~ 0.01 N(0,1) from a seed-1 stream (the factory's zero start
makes every integrand constant, so the step is measured at a nonzero point).

A step is one evaluation of loss = mean(0.5 |T|^2 - logdet) and its
coefficient gradient with inputs resident in HBM.  "fused" is one launch over
all components (mpart_kl_value_and_grad); "plain" is the reference's structure
(per-component Evaluate / LogDeterminant / CoeffGrad / LogDeterminantCoeffGrad
launches with stored Jacobians, then a row reduction).

Prints one JSON line.  --gpus must be 1.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KL fwd+adjoint points/s, fused batched vs plain GPU path @1e-10 rel err"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=500_000)
    ap.add_argument("--coeff-scale", type=float, default=0.01)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard points "
              f"across GPUs with one all-reduce of the gradient; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, banana_points
    from mpart_amd.roofline import fused_flops

    dev = torch.device("cuda:0")
    D, order, N = 8, 3, args.n
    x_np = banana_points(N, D, seed=0)
    x = torch.from_numpy(x_np).to(dev)
    tri = mt.CreateTriangular(D, D, order, mt.MapOptions())
    coeffs = args.coeff_scale * Stream(1).normal(tri.numCoeffs)
    tri.SetCoeffs(coeffs)
    stream = torch.cuda.current_stream()

    def fused():
        return tri.KLValueAndGrad(x)

    def plain():
        return tri.KLValueAndGrad(x, plain=True)

    # Warm-up (allocates the work buffers).
    for _ in range(args.warmup):
        fused()
    torch.cuda.synchronize()

    # Timed fused steps, bracketed by synchronizations; per-launch events on the
    # stream the kernels run on.
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        loss, grad = fused()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    ev_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    fused_loss = float(loss.item())
    fused_grad = grad.cpu().numpy()

    # Plain path.
    plain_ms = None
    rel_err = None
    if not args.no_plain:
        plain()
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for _ in range(args.plain_steps):
            lp, gp = plain()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - tp0) * 1e3 / args.plain_steps
        gp = gp.cpu().numpy()
        rel_err = float(max(abs(float(lp.item()) - fused_loss) / abs(fused_loss),
                            np.max(np.abs(gp - fused_grad)) / np.max(np.abs(gp))))

    # Algorithmic work from the adaptive trees of these inputs (outside the timed region).
    msets = [tri.GetMultiIndexSet(c) for c in range(tri.NumComponents())]
    _, counts = tri.EvaluateCounted(x)
    counts = counts.cpu().numpy()
    flops = fused_flops(msets, counts)
    achieved_tflops = flops / (ev_ms * 1e-3) / 1e12
    bytes_alg = 8.0 * D * N + 8.0 * (tri.numCoeffs + 1)

    out = {
        "metric": METRIC,
        "value": N / (ms_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (banana recursion, seed 0; coefficients 0.01 N(0,1), seed 1)",
        "config": {"workload": "C3: KL fwd+adjoint, 8-D triangular map, total order 3 (494 coeffs), "
                               "ProbabilistHermite/SoftPlus/AdaptiveSimpson, N=500000",
                   "N": N, "dim": D, "order": order, "num_coeffs": tri.numCoeffs},
        "fused_kernel_ms": ev_ms,
        "plain_ms_per_step": plain_ms,
        "plain_points_per_s": (N / (plain_ms * 1e-3)) if plain_ms else None,
        "speedup_fused_vs_plain": (plain_ms / ms_step) if plain_ms else None,
        "fused_vs_plain_max_rel_err": rel_err,
        "mean_integrand_evals_per_point_component": float(counts.mean()),
        "roofline": {"bound": "fp64-valu", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS, "traffic": None,
                     "algorithmic_flops_per_launch": flops, "algorithmic_bytes_per_launch": bytes_alg,
                     "hbm_GBs": bytes_alg / (ev_ms * 1e-3) / 1e9},
        "cpu_baseline": None,
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x_np, coeffs, D, order, args.cpu_seconds)

    print(json.dumps(out), flush=True)


def cpu_baseline(x_np, coeffs, D, order, budget_s):
    """The oracle (CPU restatement of the reference path) on a bounded sample."""
    from oracle import oracle as orc

    threads = min(16, os.cpu_count() or 1)
    orc.set_threads(threads)
    r = orc.CreateTriangular(D, D, order, orc.options())
    r.SetCoeffs(coeffs)
    n = 500
    sample = np.ascontiguousarray(x_np[:, :n])
    t0 = time.perf_counter()
    r.KL(sample)
    dt = time.perf_counter() - t0
    # Grow the sample to about budget_s of CPU time.
    n2 = int(min(x_np.shape[1], max(n, n * budget_s / max(dt, 1e-3) * 0.8)))
    sample = np.ascontiguousarray(x_np[:, :n2])
    t0 = time.perf_counter()
    r.KL(sample)
    dt = time.perf_counter() - t0
    return {"value": n2 / dt, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"first {n2} of the {x_np.shape[1]} points, one KL value+gradient, {dt:.1f} s"}


if __name__ == "__main__":
    main()
