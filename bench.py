"""Benchmark of the headline workload: fused forward+adjoint KL gradient.

Workload: 8-D triangular map, total order 3 per component (494 coefficients),
ProbabilistHermite / SoftPlus / adaptive Simpson (MapOptions defaults),
N = 500,000 FP64 points from the banana recursion
x1 ~ N(0,1), x_k = 0.5 x_{k-1}^2 + tanh(x_{k-2}) + 0.3 N(0,1) (x_0 := 0, seed 0),
at the factory's initial coefficients (zeros).  The recursion explodes (x8
reaches 6e48), so at any nonzero coefficient vector SoftPlus underflows for
some points and the objective is +inf; the zero start is the workload's only
finite point.  A second workload, reported beside it, uses the same map with
x ~ N(0, I_8) and coefficients 0.1 N(0,1) (seed 1), where the adaptive
quadrature subdivides.

A step is one evaluation of loss = mean(0.5 |T|^2 - logdet) and its
coefficient gradient with inputs resident in HBM and outputs preallocated.
"fused" is one launch over all components plus one column reduction
(mpart_kl_value_and_grad); "plain" is the reference's structure
(per-component Evaluate / LogDeterminant / CoeffGrad / LogDeterminantCoeffGrad
launches with stored Jacobians, then row reductions).

Prints one JSON line.  --gpus must be 1.  Before it touches the GPU the run
starts three child processes: one compiles the map's specialized module into
MPART_CODEGEN_CACHE (a temporary directory unless set), two run
`bench.py --profile` under rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE for
the line's `roofline.traffic` (--no-traffic skips them); all later processes
load the one compiled code object.

--profile runs only the headline fused steps, a 256 MiB elementwise copy (the
calibration kernel for FETCH_SIZE / WRITE_SIZE) and the active workload's
fused steps, for use under rocprofv3; --prepare-codegen only compiles the
map's specialized module;
--profile-inverse runs only three Inverse calls of the single-component
workload; --profile-many runs only the many-maps workload at dims 2..8,
orders 2..4 (Evaluate + LogDeterminant graph, then the Evaluate +
LogDeterminant + CoeffGrad + Inverse graph, each beside its per-map loop);
--profile-inverse-many runs only the batched Inverse of the many-maps
workload at the full ranges (dims 2..16, orders 2..5), once with each
schedule.
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
D, ORDER = 8, 3
NUM_COEFFS = 494  # sum over k = 1..8 of C(k + 3, 3)
CG_CHUNK_BYTES = 16 << 30  # the many-maps CoeffGrad output passes through a buffer of at most this size


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_fused(tri, x, steps, warmup):
    """Wall clock around `steps` fused launches (bracketed by synchronizations)
    and HIP events around each launch on the launch stream."""
    import torch

    stream = torch.cuda.current_stream()
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(tri.numCoeffs, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        tri.KLValueAndGradImpl(x, loss, grad)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(stream)
        tri.KLValueAndGradImpl(x, loss, grad)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    ev_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    return ms_step, ev_ms, float(loss.item()), grad.cpu().numpy()


def timed_plain(tri, x, steps):
    import torch

    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(tri.numCoeffs, dtype=torch.float64, device="cuda")
    tri.KLValueAndGradImpl(x, loss, grad, plain=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tri.KLValueAndGradImpl(x, loss, grad, plain=True)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, float(loss.item()), grad.cpu().numpy()


def measure(tri, x, steps, warmup, plain_steps, do_plain):
    from mpart_amd.roofline import algo_flops, fused_flops

    N = x.shape[1]
    ms_step, ev_ms, fl, fg = timed_fused(tri, x, steps, warmup)
    log(f"fused {ms_step:.3f} ms/step (events {ev_ms:.3f} ms), loss {fl:.6g}")
    res = {"ms_per_step": ms_step, "event_ms": ev_ms, "points_per_s": N / (ms_step * 1e-3), "loss": fl}
    if do_plain:
        pms, lp, gp = timed_plain(tri, x, plain_steps)
        res["plain_ms_per_step"] = pms
        res["plain_points_per_s"] = N / (pms * 1e-3)
        res["speedup_fused_vs_plain"] = pms / ms_step
        res["fused_vs_plain_max_rel_err"] = float(max(abs(lp - fl) / abs(fl),
                                                      np.max(np.abs(gp - fg)) / np.max(np.abs(gp))))
        log(f"plain {pms:.3f} ms/step")
        # The same fused pass through the precompiled generic kernel.
        tri.SetCodegen(False)
        gms, _, _, _ = timed_fused(tri, x, plain_steps, 1)
        tri.SetCodegen(True)
        res["generic_fused_ms_per_step"] = gms
        log(f"generic fused {gms:.3f} ms/step")
    msets = [tri.GetMultiIndexSet(c) for c in range(tri.NumComponents())]
    _, counts = tri.EvaluateCounted(x)
    counts = counts.cpu().numpy()
    flops = algo_flops(msets, counts)
    regrouped = fused_flops(msets, counts)
    res["mean_integrand_evals"] = float(counts.mean())
    res["max_integrand_evals"] = int(counts.max())
    res["algorithmic_flops"] = flops
    res["algorithmic_flops_per_point"] = flops / N
    res["achieved_tflops"] = flops / (ev_ms * 1e-3) / 1e12
    res["regrouped_model_flops"] = regrouped
    res["regrouped_model_flops_per_point"] = regrouped / N
    res["regrouped_model_tflops"] = regrouped / (ev_ms * 1e-3) / 1e12
    return res, fg


def oracle_check(x_np, coeffs, fused_grad_fn, n=2000):
    """Max relative error of the GPU fused result against the oracle on the first n points."""
    from oracle import oracle as orc

    r = orc.CreateTriangular(D, D, ORDER, orc.options())
    r.SetCoeffs(coeffs)
    xs = np.ascontiguousarray(x_np[:, :n])
    lr, gr = r.KL(xs)
    lg, gg = fused_grad_fn(xs)
    return float(max(abs(lg - lr) / abs(lr), np.max(np.abs(gg - gr)) / np.max(np.abs(gr))))


def many_maps(steps, dmax=8, pmax=4, nmax=16384, coeff_grad=True, inverse=True):
    """256 independent triangular maps (dims 2..dmax, orders 2..pmax,
    4096..nmax N(0, I) points each, coefficients 0.1 N(0,1); seed-21 stream).

    Evaluate + LogDeterminant of every map as one batched launch replayed from
    a HIP graph, against a host loop of per-map launches (generic kernels).
    Then the whole sequence in one graph: Evaluate + LogDeterminant, CoeffGrad
    (sens from a seed-22 stream; `coeff_grad`) and Inverse of the Evaluate
    output (`inverse`), against the per-map loop of the same calls.  The
    Inverse recovers the points, so max |Inverse(T(x)) - x| is a check at the
    full size."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream

    s = Stream(21)
    nm = 256
    u = s.uniform(3 * nm).reshape(nm, 3)
    maps, xs = [], []
    for i in range(nm):
        d = 2 + int(u[i, 0] * (dmax - 1))
        p = 2 + int(u[i, 1] * (pmax - 1))
        n = 4096 + int(u[i, 2] * (nmax - 4096))
        m = mt.CreateTriangular(d, d, p, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        m.SetCodegen(False)
        maps.append(m)
        xs.append(torch.from_numpy(s.points(d, n)).cuda())
    npts = sum(x.shape[1] for x in xs)
    ys = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    lds = [torch.empty(x.shape[1], dtype=torch.float64, device="cuda") for x in xs]
    b = mt.MapBatch(maps)

    def capture(fn):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fn()  # eager first call: tunes the schedule, uploads the arguments
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        g.replay()
        torch.cuda.synchronize()
        return g

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    g = capture(lambda: b.EvaluateImpl(xs, ys, lds))
    tb = timed(g.replay, steps)
    y_b = [t.clone() for t in ys]

    def loop_eval():
        for m, x, y, ld in zip(maps, xs, ys, lds):
            m.EvaluateImpl(x, y)
            m.LogDeterminantImpl(x, ld)

    loop_eval()
    tl = timed(loop_eval, steps)
    same = all(torch.equal(a, c) for a, c in zip(y_b, ys))
    maxabs = max((a - c).abs().max().item() for a, c in zip(y_b, ys))
    log(f"many maps (dims 2..{dmax}, orders 2..{pmax}, N <= {nmax}): batched graph {tb:.3f} ms, "
        f"per-map loop {tl:.3f} ms, identical outputs {same}, max abs diff {maxabs:.2e}")
    res = {"maps": nm, "points": int(npts), "num_coeffs": int(sum(m.numCoeffs for m in maps)),
           "batched_graph_ms": tb, "per_map_loop_ms": tl,
           "batched_points_per_s": npts / (tb * 1e-3), "speedup_batched_vs_loop": tl / tb,
           "batched_equals_per_map_bitwise": bool(same), "max_abs_diff_batched_vs_loop": maxabs,
           "batched_schedule": b.EvalSchedule()[0], "batched_trial_ms_schedules_1_4": b.EvalSchedule()[1],
           "workload": f"Evaluate + LogDeterminant of 256 triangular maps, dims 2..{dmax}, orders 2..{pmax}, "
                       f"4096..{nmax} N(0,I) points each"}
    if not (coeff_grad or inverse):
        return res

    # the whole sequence in one graph
    cg_bytes = 8 * sum(m.numCoeffs * x.shape[1] for m, x in zip(maps, xs))
    cg_calls = []  # (batch, points, sens, outputs) of each CoeffGrad call
    if coeff_grad:
        ss = Stream(22)
        sens = [torch.from_numpy(ss.normal(m.outputDim * x.shape[1]).reshape(m.outputDim, x.shape[1])).cuda()
                for m, x in zip(maps, xs)]
        nchunks = max(1, -(-cg_bytes // CG_CHUNK_BYTES))
        if nchunks == 1:
            cgs = [torch.empty((m.numCoeffs, x.shape[1]), dtype=torch.float64, device="cuda")
                   for m, x in zip(maps, xs)]
            cg_calls.append((b, xs, sens, cgs))
        else:
            # Outputs larger than one GPU holds: the points of every map in
            # nchunks slices, each slice's CoeffGrad written into one reused
            # buffer (a batch per slice: a captured batch's arguments are fixed).
            sizes = [[(x.shape[1] * (k + 1)) // nchunks - (x.shape[1] * k) // nchunks for x in xs]
                     for k in range(nchunks)]
            buf = torch.empty(max(sum(m.numCoeffs * n for m, n in zip(maps, sz)) for sz in sizes),
                              dtype=torch.float64, device="cuda")
            for k in range(nchunks):
                xk, sk, ok, off = [], [], [], 0
                for m, x, sv, n in zip(maps, xs, sens, sizes[k]):
                    a = (x.shape[1] * k) // nchunks
                    xk.append(x[:, a:a + n])
                    sk.append(sv[:, a:a + n])
                    ok.append(buf[off:off + m.numCoeffs * n].view(m.numCoeffs, n))
                    off += m.numCoeffs * n
                cg_calls.append((mt.MapBatch(maps), xk, sk, ok))
            cgs = cg_calls[-1][3]
    x1s = [x[:0] for x in xs]  # square triangular maps: no conditioning rows
    xis = [torch.empty_like(x) for x in xs]

    def seq():
        b.EvaluateImpl(xs, ys, lds)
        for bk, xk, sk, ok in cg_calls:
            bk.CoeffGradImpl(xk, sk, ok)
        if inverse:
            b.InverseImpl(x1s, ys, xis)

    gs = capture(seq)
    reps = max(1, steps // 2)
    ts = timed(gs.replay, reps)
    out_b = [[t.clone() for t in ts_] for ts_ in ((cgs if coeff_grad else []), (xis if inverse else []))]

    def loop_seq():
        for i, (m, x) in enumerate(zip(maps, xs)):
            m.EvaluateImpl(x, ys[i])
            m.LogDeterminantImpl(x, lds[i])
        # slice by slice, as the graph does, so that the shared buffer ends
        # with the last slice's outputs
        for _, xk, sk, ok in cg_calls:
            for i, m in enumerate(maps):
                m.CoeffGradImpl(xk[i], sk[i], ok[i])
        if inverse:
            for i, m in enumerate(maps):
                m.InverseImpl(x1s[i], ys[i], xis[i])

    loop_seq()
    tls = timed(loop_seq, 1)
    def same_bits(a, c):
        # NaN-aware: an Inverse whose bracket search fails returns NaN, as the reference's does
        return bool(((a == c) | (torch.isnan(a) & torch.isnan(c))).all().item())

    outs_now = (cgs if coeff_grad else []) + (xis if inverse else [])
    seq_same = all(same_bits(a, c) for a, c in zip(out_b[0] + out_b[1], outs_now))
    names = ["Evaluate + LogDeterminant"] + (["CoeffGrad"] if coeff_grad else []) + (["Inverse"] if inverse else [])
    res["sequence"] = {"methods": " + ".join(names), "batched_graph_ms": ts, "per_map_loop_ms": tls,
                       "speedup_batched_vs_loop": tls / ts, "batched_equals_per_map_bitwise": seq_same,
                       "coeff_grad_output_bytes": cg_bytes if coeff_grad else None,
                       "coeff_grad_launches": len(cg_calls),
                       "coeff_grad_note": ("each map's points in slices whose outputs share one reused buffer; the "
                                           "bitwise comparison covers the last slice" if len(cg_calls) > 1 else None)}
    if coeff_grad:
        res["sequence"]["coeff_grad_max_rel_diff_batched_vs_loop"] = max(
            ((a - c).abs().max() / c.abs().max().clamp_min(1e-300)).item() for a, c in zip(out_b[0], cgs))
    if inverse:
        # per point: recovered (all coordinates finite) or not; round-trip error
        # over recovered points.  Where T is nearly flat in its last input the
        # solver meets the y tolerance far from x, and the later components
        # inherit that error, so the error is reported, not bounded.
        fin = [torch.isfinite(a).all(0) for a in out_b[1]]
        ex = [(a - x).abs().amax(0)[f] for a, x, f in zip(out_b[1], xs, fin)]
        res["sequence"]["inverse_points_not_finite"] = int(sum((~f).sum().item() for f in fin))
        res["sequence"]["inverse_roundtrip_max_abs_err_finite"] = max(e.max().item() if e.numel() else 0.0 for e in ex)
        res["sequence"]["inverse_points_err_above_1e-6"] = int(sum((e > 1e-6).sum().item() for e in ex))
        res["sequence"]["inverse_points_err_above_1e-3"] = int(sum((e > 1e-3).sum().item() for e in ex))
        # maps where some call met an infinite g (the reference would have
        # failed there with its runtime error); their solves end with NaN
        res["sequence"]["maps_with_infinite_g"] = _maps_with_infinite_g(maps)
        res["sequence"]["inverse_handover"] = b.InverseHandoverStats()
        sched, trial = b.InverseSchedule()
        res["sequence"]["inverse_schedule"] = {"choice": sched, "trial_ms_schedules_1_2": trial,
                                               "meaning": "1: per-lane solve scheduling, one pass per component "
                                                          "index; 2: a lane per point, all components in one launch"}
    log(f"many maps sequence ({res['sequence']['methods']}): batched graph {ts:.3f} ms, per-map loop {tls:.3f} ms, "
        f"identical {seq_same}; {json.dumps({k: v for k, v in res['sequence'].items() if k.startswith(('inverse', 'coeff'))})}")
    return res


def profile_inverse_many_mode():
    """Batched Inverse of the many-maps workload at the full ranges (the maps
    and points of many_maps(dmax=16, pmax=5, nmax=65536), targets T(x)), once
    with each schedule (1: per-lane solve scheduling, 2: a lane per point)."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream

    s = Stream(21)
    nm, dmax, pmax, nmax = 256, 16, 5, 65536
    u = s.uniform(3 * nm).reshape(nm, 3)
    maps, xs = [], []
    for i in range(nm):
        d = 2 + int(u[i, 0] * (dmax - 1))
        p = 2 + int(u[i, 1] * (pmax - 1))
        n = 4096 + int(u[i, 2] * (nmax - 4096))
        m = mt.CreateTriangular(d, d, p, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        m.SetCodegen(False)
        maps.append(m)
        xs.append(torch.from_numpy(s.points(d, n)).cuda())
    b = mt.MapBatch(maps)
    ys = b.Evaluate(xs)
    x1s = [x[:0] for x in xs]
    xis = [torch.empty_like(x) for x in xs]
    out = {}
    for sched in (1, 2):
        b.SetInverseSchedule(sched)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.InverseImpl(x1s, ys, xis)
        torch.cuda.synchronize()
        out[f"schedule_{sched}_ms"] = (time.perf_counter() - t0) * 1e3
        out[f"schedule_{sched}_nan_roots"] = int(sum(int(torch.isnan(x).sum().item()) for x in xis))
        out[f"schedule_{sched}_maps_with_infinite_g"] = _maps_with_infinite_g(maps)
        if sched == 1:
            out["schedule_1_handover"] = b.InverseHandoverStats()
        log(f"inverse schedule {sched}: {out[f'schedule_{sched}_ms']:.1f} ms")
    print(json.dumps({"profile": "inverse_many_maps_full_ranges", **out}), flush=True)


def _maps_with_infinite_g(maps):
    """How many maps met an infinite g at a quadrature node since their last
    check (the reference's runtime-error condition); clears the records."""
    n = 0
    for m in maps:
        try:
            m.CheckIntegrand()
        except RuntimeError:
            n += 1
    return n


def single_component_setup():
    """The single-component workload's map, points, targets (see below)."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream

    opts = mt.MapOptions(quadType=mt.QuadTypes.AdaptiveClenshawCurtis, quadPts=5, quadAbsTol=1e-8, quadRelTol=1e-8)
    g = mt.CreateTotalOrderComponent(6, 4, opts)
    s = Stream(0)
    N = 1_000_000
    x = torch.from_numpy(s.points(6, N)).cuda()
    g.SetCoeffs(0.05 * s.normal(g.numCoeffs))
    r = torch.from_numpy(s.normal(N)[None, :].copy()).cuda()
    return g, x, r, N


def profile_inverse_mode(args):
    """Three Inverse calls of the single-component workload (for rocprofv3)."""
    import torch

    g, x, r, N = single_component_setup()
    xi = torch.empty((1, N), dtype=torch.float64, device="cuda")
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.InverseImpl(x, r, xi)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    print(json.dumps({"profile": "inverse_c2_random_targets", "ms": times,
                      "nan": int(torch.isnan(xi).sum().item()), "handover": g.InverseHandoverStats()}), flush=True)


def single_component_workload(steps):
    """One component, dim 6, total order 4 (210 terms), ProbabilistHermite /
    SoftPlus / adaptive Clenshaw-Curtis (quadPts 5, tolerances 1e-8);
    N = 1,000,000 points x ~ N(0, I_6), coefficients ~ N(0, 0.05^2), targets
    r ~ N(0, 1) from one seed-0 stream.  Times Evaluate, LogDeterminant,
    CoeffGrad (sens = 1, 210 x N output) and Inverse (x_{1:5} fixed, x_6 as
    the initial guess) with preallocated outputs."""
    import torch

    g, x, r, N = single_component_setup()
    g.PrepareCodegen()
    y = torch.empty((1, N), dtype=torch.float64, device="cuda")
    ld = torch.empty(N, dtype=torch.float64, device="cuda")
    sens = torch.ones((1, N), dtype=torch.float64, device="cuda")
    cg = torch.empty((g.numCoeffs, N), dtype=torch.float64, device="cuda")
    xi = torch.empty((1, N), dtype=torch.float64, device="cuda")
    calls = {
        "evaluate": lambda: g.EvaluateImpl(x, y),
        "log_determinant": lambda: g.LogDeterminantImpl(x, ld),
        "coeff_grad": lambda: g.CoeffGradImpl(x, sens, cg),
        "inverse": lambda: g.InverseImpl(x, r, xi),
    }
    out = {"workload": "one component, dim 6, total order 4 (210 terms), AdaptiveClenshawCurtis tol 1e-8, "
                       "N=1000000", "N": N}
    for name, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out[name + "_ms"] = ms
        out[name + "_points_per_s"] = N / (ms * 1e-3)
        log(f"single component {name}: {ms:.3f} ms")
    # the Inverse's tail: solves handed over to a wavefront each (last call)
    out["inverse_handover"] = g.InverseHandoverStats()
    del cg
    return out


def input_grad_workload(tri, x, steps):
    """Point gradient of the KL objective, grad_x [0.5 |T|^2 - logdet] / N, on the
    headline map and the given points: the fused launch (KLInputGradImpl)
    against the unfused sequence the reference's torch wrapper runs (Evaluate,
    Gradient with sens = T, LogDeterminantInputGrad, then the difference);
    and the loss with both gradients from one pass (KLValueAndGradsImpl)
    against the loss with the coefficient gradient only."""
    import torch

    D_, N = x.shape
    y = torch.empty((tri.outputDim, N), dtype=torch.float64, device="cuda")
    g1 = torch.empty((D_, N), dtype=torch.float64, device="cuda")
    g2 = torch.empty((D_, N), dtype=torch.float64, device="cuda")
    gf = torch.empty((D_, N), dtype=torch.float64, device="cuda")

    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(tri.numCoeffs, dtype=torch.float64, device="cuda")
    g1p = torch.empty((D_, N), dtype=torch.float64, device="cuda")

    def fused():
        tri.KLInputGradImpl(x, gf, 1.0 / N)

    def coeff_only():
        tri.KLValueAndGradImpl(x, loss, grad)

    def one_pass():
        tri.KLValueAndGradsImpl(x, loss, grad, g1p)

    def unfused():
        tri.EvaluateImpl(x, y)
        tri.GradientImpl(x, y, g1)
        tri.LogDeterminantInputGradImpl(x, g2)
        torch.sub(g1, g2, out=g1).mul_(1.0 / N)

    out = {"N": N}
    for name, fn in (("fused", fused), ("unfused", unfused), ("kl_coeff_grad", coeff_only),
                     ("kl_both_grads_one_pass", one_pass)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        out[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / steps
        log(f"KL input gradient {name}: {out[name + '_ms']:.3f} ms")
    out["speedup"] = out["unfused_ms"] / out["fused_ms"]
    out["max_rel_diff"] = ((gf - g1).abs().max() / g1.abs().max()).item()
    out["one_pass_max_rel_diff"] = ((g1p - g1).abs().max() / g1.abs().max()).item()
    return out


def training_workload(tri, x, steps=50):
    """KL training of the 8-D order-3 map with torch.optim.LBFGS (`steps`
    iterations, strong-Wolfe line search, from the factory's zero
    coefficients) on the points x: through the fused objective (one fused
    pass per evaluation, mpart_amd.torch.kl_objective) and through the
    reference's torch structure (Evaluate + LogDeterminant forward, the
    coefficient VJP backward, mpart_amd.torch.kl_loss).  Each path runs once
    untimed (first calls compile, tune and size buffers), then timed; the
    time inside objective evaluations (forward + backward, synchronized) is
    reported beside the wall time, which adds the optimizer's host work."""
    import torch

    from mpart_amd.torch import kl_loss, kl_objective

    zeros = torch.zeros(tri.numCoeffs, dtype=torch.float64, device="cuda")

    def run(fn):
        coeffs = torch.nn.Parameter(zeros.clone())
        opt = torch.optim.LBFGS([coeffs], lr=1.0, max_iter=steps, max_eval=int(steps * 1.25) + 1, history_size=20,
                                tolerance_grad=1e-12, tolerance_change=1e-14, line_search_fn="strong_wolfe")
        hist, evals = [], []

        def closure():
            opt.zero_grad()
            t0 = time.perf_counter()
            loss = fn(tri, x, coeffs)
            loss.backward()
            hist.append(float(loss.detach()))  # synchronizes
            evals.append(time.perf_counter() - t0)
            return loss

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.step(closure)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, hist, sum(evals)

    out = {"steps": steps, "N": int(x.shape[1])}
    for name, fn in (("fused", kl_objective), ("unfused", kl_loss)):
        run(fn)  # untimed
        wall, hist, ev = run(fn)
        out[f"{name}_s"] = wall
        out[f"{name}_objective_s"] = ev
        out[f"{name}_evaluations"] = len(hist)
        out[f"{name}_final_loss"] = hist[-1]
        out[f"_{name}_hist"] = hist
    hf, hu = out.pop("_fused_hist"), out.pop("_unfused_hist")
    out["speedup_fused_vs_unfused_wall"] = out["unfused_s"] / out["fused_s"]
    out["speedup_fused_vs_unfused_objective"] = out["unfused_objective_s"] / out["fused_objective_s"]
    n = min(len(hf), len(hu))
    out["objective_history_max_rel_diff"] = max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(hf[:n], hu[:n]))
    tri.SetCoeffs(zeros)
    log(f"training ({steps} L-BFGS steps, N={x.shape[1]}): fused {out['fused_s'] * 1e3:.1f} ms "
        f"(objective {out['fused_objective_s'] * 1e3:.1f} ms, {out['fused_evaluations']} evaluations), unfused "
        f"{out['unfused_s'] * 1e3:.1f} ms (objective {out['unfused_objective_s'] * 1e3:.1f} ms), "
        f"history max rel diff {out['objective_history_max_rel_diff']:.1e}")
    return out


def uneven_map_workload(steps, plain_n=20_000):
    """32-D triangular map, component k of total order max(1, 6 - k // 6)
    (7 to 7 315 terms, 62 572 coefficients), HermiteFunctions / Exp / adaptive
    Simpson (tolerances 1e-6); N = 200 000 points with Student-t(5)
    coordinates (seed 3), coefficients 0.01 N(0, 1) (seed 5).  Above the
    specialization limit, so the generic kernels run with the autotuned
    point schedule (lanes per point).  Times Evaluate, LogDeterminant and the
    fused KL value + gradient; the plain KL path (per-component launches with
    the two K x N Jacobians stored) on the first plain_n points (20 000 by default)."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, student_t_points

    orders = [max(1, 6 - k // 6) for k in range(32)]
    o = mt.MapOptions(basisType=mt.BasisTypes.HermiteFunctions, posFuncType=mt.PosFuncTypes.Exp)
    g = mt.TriangularMap([mt.CreateTotalOrderComponent(k + 1, orders[k], o) for k in range(32)])
    g.SetCoeffs(0.01 * Stream(5).normal(g.numCoeffs))
    N = 200_000
    x = torch.from_numpy(student_t_points(32, N, nu=5, seed=3)).cuda()
    y = torch.empty((32, N), dtype=torch.float64, device="cuda")
    ld = torch.empty(N, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
    xs = x[:, :plain_n].contiguous()
    calls = {
        "evaluate": lambda: g.EvaluateImpl(x, y),
        "log_determinant": lambda: g.LogDeterminantImpl(x, ld),
        "kl_fused": lambda: g.KLValueAndGradImpl(x, loss, grad),
        f"kl_fused_{plain_n}": lambda: g.KLValueAndGradImpl(xs, loss, grad),
        f"kl_plain_{plain_n}": lambda: g.KLValueAndGradImpl(xs, loss, grad, plain=True),
    }
    out = {"workload": "32-D triangular, orders max(1, 6 - k // 6), 62572 coefficients, HermiteFunctions/Exp/"
                       "AdaptiveSimpson, N=200000 Student-t(5)", "N": N, "num_coeffs": g.numCoeffs}
    for name, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out[name + "_ms"] = ms
        log(f"uneven 32-D map {name}: {ms:.2f} ms")
    out["kl_fused_points_per_s"] = N / (out["kl_fused_ms"] * 1e-3)
    out[f"fused_vs_plain_{plain_n}"] = out[f"kl_plain_{plain_n}_ms"] / out[f"kl_fused_{plain_n}_ms"]
    # Roofline of the fused pass: the reference-algorithm flop count of these
    # points' adaptive trees (mpart_amd/roofline.py algo_flops) over its wall time.
    from mpart_amd.roofline import algo_flops

    msets = [g.GetMultiIndexSet(c) for c in range(g.NumComponents())]
    _, counts = g.EvaluateCounted(x)
    counts = counts.cpu().numpy()
    flops = algo_flops(msets, counts, exp_bijector=True)
    out["mean_integrand_evals"] = float(counts.mean())
    out["max_integrand_evals"] = int(counts.max())
    out["kl_fused_algorithmic_flops"] = flops
    out["kl_fused_achieved_tflops"] = flops / (out["kl_fused_ms"] * 1e-3) / 1e12
    out["kl_fused_fp64_frac"] = out["kl_fused_achieved_tflops"] / 78.6
    lanes, trials = g.EvalSchedule()
    out["lanes_per_point"] = {"evaluate": lanes[0], "log_determinant": lanes[1], "fused": lanes[2],
                              "trial_ms_lanes_1_4_8_64_split": {"evaluate": trials[0],
                                                                "log_determinant": trials[1],
                                                                "fused": trials[2]},
                              "meaning": "lanes per point; 2: components split over the wavefronts of a block"}
    return out


def profile_mode(args):
    """For rocprofv3: the headline fused steps (autotuning trials, warm-up,
    `steps` timed steps), then the calibration kernel (torch.mul over 256 MiB:
    reads 256 MiB, writes 256 MiB), then `steps` fused steps of the active
    workload (N(0, I) points, 0.1 N(0, 1) coefficients, seed-1 stream) on the
    same map and schedule.  Prints one JSON line with the schedule, the grid
    and the fused kernel's resources in this process."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, banana_points

    tri = mt.CreateTriangular(D, D, ORDER, mt.MapOptions())
    tri.PrepareCodegen()
    x = torch.from_numpy(banana_points(args.n, D, seed=0)).cuda()
    tri.SetCoeffs(np.zeros(tri.numCoeffs))
    ms, ev, fl, _ = timed_fused(tri, x, args.steps, args.warmup)
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)  # reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    del a, b
    s = Stream(1)
    x2 = torch.from_numpy(s.points(D, args.n)).cuda()
    tri.SetCoeffs(0.1 * s.normal(tri.numCoeffs))
    ms2, ev2, fl2, _ = timed_fused(tri, x2, args.steps, 0)
    choice, trials = tri.FusedSchedule()
    print(json.dumps({"profile": "kl_fused_c3", "headline_event_ms": ev, "active_event_ms": ev2, "loss": fl,
                      "active_loss": fl2, "fused_schedule": choice, "trial_ms": trials,
                      "kernel": kernel_info(tri, choice),
                      "calibration": "torch.mul over 256 MiB (reads 262144 KiB, writes 262144 KiB)"}), flush=True)


def kernel_info(tri, choice):
    """Resources of the fused kernel the schedule runs, as this process compiled it."""
    name = "fused_w3" if (choice is not None and choice >= 0 and choice & 1) else "fused"
    try:
        v, sc, lds = tri.KernelInfo(name)
    except Exception as e:  # not specialized
        return {"kernel": name, "error": str(e)}
    return {"kernel": f"mpart_gen_{name}", "vgprs": v, "scratch_bytes_per_lane": sc, "lds_bytes_static": lds}


def _pmc_pass(counter, steps, n, outdir, timeout_s):
    """One rocprofv3 --pmc pass over `bench.py --profile`; returns
    ([(dispatch id, kernel name, grid blocks, value)], the child's JSON line)."""
    import csv
    import glob
    import shutil
    import subprocess

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    d = os.path.join(outdir, counter)
    cmd = [exe, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--",
           sys.executable, os.path.abspath(__file__), "--profile", "--steps", str(steps), "--warmup", "1",
           "--n", str(n)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout_s)
    child = None
    for line in r.stdout.splitlines():
        if line.startswith("{") and '"profile"' in line:
            child = json.loads(line)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file (rc {r.returncode}): {r.stderr[-800:]}")
    per = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            key = int(row["Dispatch_Id"])
            grid = int(row.get("Grid_Size", 0) or 0)
            wg = int(row.get("Workgroup_Size", 1) or 1)
            ent = per.setdefault(key, [row["Kernel_Name"], grid // max(wg, 1), 0.0])
            ent[2] += float(row["Counter_Value"])
    return [(k, v[0], v[1], v[2]) for k, v in sorted(per.items())], child


def measure_traffic(steps, n, K, timeout_s=240):
    """HBM bytes per fused launch of the headline and the active workload:
    rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE (separate child runs of
    `bench.py --profile`, started before this process touches the GPU), each
    counter scaled by what it reads for the 256 MiB calibration copy.  Values
    are per dispatch (summed over counter instances), averaged over the last
    `steps` fused dispatches of each workload."""
    import tempfile

    res = {"method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over `bench.py --profile`, scaled by a "
                     "256 MiB torch.mul calibration copy"}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as outdir:
        for counter, side in (("FETCH_SIZE", "read"), ("WRITE_SIZE", "write")):
            rows, child = _pmc_pass(counter, steps, n, outdir, timeout_s)
            calib = [r for r in rows if "elementwise" in r[1]]
            if not calib:
                raise RuntimeError("no calibration dispatch")
            cal = max(calib, key=lambda r: r[3])
            factor = 256.0 * 1024 * 1024 / (cal[3] * 1024.0)
            fused = [r for r in rows if "mpart_gen_fused" in r[1]]
            for wl, sel in (("headline", [r for r in fused if r[0] < cal[0]]),
                            ("active", [r for r in fused if r[0] > cal[0]])):
                sel = sel[-steps:]
                w = res.setdefault(wl, {})
                w[f"{side}_bytes"] = float(np.mean([r[3] for r in sel])) * 1024.0 * factor
                w["blocks"] = sel[-1][2]
                w[f"{counter}_KiB"] = float(np.mean([r[3] for r in sel]))
            res[f"{side}_factor"] = factor
            res["child"] = child
    for wl in ("headline", "active"):
        w = res[wl]
        w["bytes_per_launch"] = w["read_bytes"] + w["write_bytes"]
        w["algorithmic_bytes"] = 8.0 * D * n + 8.0 * K + 8.0 * (K + 1) * w["blocks"]
    return res


def profile_grads_mode(args):
    """For rocprofv3: on the 8-D order-3 map with N(0, I) data (seed 1) and
    0.1 N(0, 1) coefficients, N = 500 000: three KL point-gradient launches
    (KLInputGradImpl) and three one-pass loss + both gradients
    (KLValueAndGradsImpl); then on the 32-D uneven map (orders
    max(1, 6 - k // 6), HermiteFunctions / Exp, Student-t(5), 20 000 points)
    Evaluate, LogDeterminant and the fused KL once each with 1 and with 8
    lanes per point."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, student_t_points

    tri = mt.CreateTriangular(D, D, ORDER, mt.MapOptions())
    s = Stream(1)
    x = torch.from_numpy(s.points(D, args.n)).cuda()
    tri.SetCoeffs(0.1 * s.normal(tri.numCoeffs))
    gx = torch.empty((D, args.n), dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(tri.numCoeffs, dtype=torch.float64, device="cuda")
    tri.KLInputGradImpl(x, gx, 1.0 / args.n)
    tri.KLValueAndGradsImpl(x, loss, grad, gx)
    torch.cuda.synchronize()
    for _ in range(3):
        tri.KLInputGradImpl(x, gx, 1.0 / args.n)
    for _ in range(3):
        tri.KLValueAndGradsImpl(x, loss, grad, gx)
    torch.cuda.synchronize()
    orders = [max(1, 6 - k // 6) for k in range(32)]
    o = mt.MapOptions(basisType=mt.BasisTypes.HermiteFunctions, posFuncType=mt.PosFuncTypes.Exp)
    g = mt.TriangularMap([mt.CreateTotalOrderComponent(k + 1, orders[k], o) for k in range(32)])
    g.SetCoeffs(0.01 * Stream(5).normal(g.numCoeffs))
    n = 20_000
    xu = torch.from_numpy(student_t_points(32, n, nu=5, seed=3)).cuda()
    y = torch.empty((32, n), dtype=torch.float64, device="cuda")
    ld = torch.empty(n, dtype=torch.float64, device="cuda")
    lossu = torch.zeros(1, dtype=torch.float64, device="cuda")
    gradu = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
    for lanes in (1, 8):
        g.SetPointGroup(lanes)
        g.EvaluateImpl(xu, y)
        g.LogDeterminantImpl(xu, ld)
        g.KLValueAndGradImpl(xu, lossu, gradu)
    torch.cuda.synchronize()
    print(json.dumps({"profile": "kl_input_grad_c3_normal + uneven_map_group_kernels", "loss": float(loss.item())}),
          flush=True)


def profile_uneven_mode(args):
    """For rocprofv3: the 32-D uneven map (orders max(1, 6 - k // 6),
    HermiteFunctions / Exp, Student-t(5) points, seed 3, N = 200 000,
    coefficients 0.01 N(0, 1), seed 5): Evaluate, LogDeterminant, then the
    fused KL value + gradient, each once untimed (the schedule tuner times
    its candidates there: lanes per point 1, 4, 8, 64 and the components split
    over the wavefronts of a block, in two register budgets) and then `steps`
    times on the schedule it picked."""
    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, student_t_points

    orders = [max(1, 6 - k // 6) for k in range(32)]
    o = mt.MapOptions(basisType=mt.BasisTypes.HermiteFunctions, posFuncType=mt.PosFuncTypes.Exp)
    g = mt.TriangularMap([mt.CreateTotalOrderComponent(k + 1, orders[k], o) for k in range(32)])
    g.SetCoeffs(0.01 * Stream(5).normal(g.numCoeffs))
    n = 200_000
    x = torch.from_numpy(student_t_points(32, n, nu=5, seed=3)).cuda()
    y = torch.empty((32, n), dtype=torch.float64, device="cuda")
    ld = torch.empty(n, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
    for fn in (lambda: g.EvaluateImpl(x, y), lambda: g.LogDeterminantImpl(x, ld),
               lambda: g.KLValueAndGradImpl(x, loss, grad)):
        fn()
        torch.cuda.synchronize()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
    lanes, trials = g.EvalSchedule()
    print(json.dumps({"profile": "uneven_map_split_kernels", "loss": float(loss.item()), "lanes": lanes,
                      "fused_trial_ms": trials[2]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=500_000)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-second", action="store_true")
    ap.add_argument("--no-many", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-input-grad", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--prepare-codegen", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--profile-inverse", action="store_true")
    ap.add_argument("--profile-grads", action="store_true")
    ap.add_argument("--profile-many", action="store_true")
    ap.add_argument("--profile-inverse-many", action="store_true")
    ap.add_argument("--profile-uneven", action="store_true")
    # points of the uneven map's plain (materialized Jacobian) comparison:
    # 2 x 62 572 x N doubles of Jacobians, 1 MB per point
    ap.add_argument("--uneven-plain-n", type=int, default=20_000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on a single GPU "
              f"(the path would shard points across GPUs with one all-reduce of the gradient).", file=sys.stderr)
        sys.exit(2)
    if args.profile:
        profile_mode(args)
        return
    if args.prepare_codegen:
        import mpart_amd as mt

        tri = mt.CreateTriangular(D, D, ORDER, mt.MapOptions())
        spec, secs, err = tri.PrepareCodegen()
        print(json.dumps({"specialized": spec, "compile_s": secs, "error": err,
                          "cache": os.environ.get("MPART_CODEGEN_CACHE", "")}), flush=True)
        return
    if args.profile_inverse:
        profile_inverse_mode(args)
        return
    if args.profile_grads:
        profile_grads_mode(args)
        return
    if args.profile_inverse_many:
        profile_inverse_many_mode()
        return
    if args.profile_uneven:
        profile_uneven_mode(args)
        return
    if args.profile_many:
        print(json.dumps({"profile": "many_maps_sequence", **many_maps(args.steps)}), flush=True)
        return

    # One compile of the specialized module for every process of this run:
    # a child process (started, like the profiler passes, before this process
    # touches the GPU) compiles it into MPART_CODEGEN_CACHE, and the profiled
    # children and this process load that code object, so the counters and
    # the timings below describe one binary.
    own_cache = None
    if not os.environ.get("MPART_CODEGEN_CACHE"):
        import tempfile

        own_cache = tempfile.mkdtemp(prefix="mpart_codegen_", dir=os.environ.get("TMPDIR", "/tmp"))
        os.environ["MPART_CODEGEN_CACHE"] = own_cache
    prep = None
    try:
        import subprocess

        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--prepare-codegen"], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=300)
        for line in r.stdout.splitlines():
            if line.startswith("{"):
                prep = json.loads(line)
    except Exception as e:
        log(f"codegen preparation failed: {e}")

    # HBM traffic of the fused kernel, from rocprofv3 counter passes run as
    # child processes before this process touches the GPU.
    traffic, traffic_err = None, None
    if not args.no_traffic:
        log("rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE)")
        try:
            traffic = measure_traffic(3, args.n, NUM_COEFFS)
        except Exception as e:  # profiler absent or failed: the line reports null traffic
            traffic_err = f"{type(e).__name__}: {e}"[:400]
            log(f"traffic measurement failed: {traffic_err}")

    import torch

    import mpart_amd as mt
    from mpart_amd.data import Stream, banana_points

    N = args.n
    tri = mt.CreateTriangular(D, D, ORDER, mt.MapOptions())
    spec, compile_s, gen_err = tri.PrepareCodegen()
    log(f"specialized kernels: {spec} ({compile_s:.2f} s) {gen_err}")

    # Headline: banana data, zero coefficients.
    x_np = banana_points(N, D, seed=0)
    x = torch.from_numpy(x_np).cuda()
    c0 = np.zeros(tri.numCoeffs)
    tri.SetCoeffs(c0)
    log(f"headline: N={N}, coeffs={tri.numCoeffs}")
    head, _ = measure(tri, x, args.steps, args.warmup, args.plain_steps, not args.no_plain)

    def gpu_kl(xs):
        l, g = tri.KLValueAndGrad(torch.from_numpy(xs).cuda())
        return float(l.item()), g.cpu().numpy()

    head["oracle_max_rel_err_2000pts"] = oracle_check(x_np, c0, gpu_kl)
    choice, trials = tri.FusedSchedule()
    fused_schedule = {"choice": choice, "trial_ms": trials,
                      "meaning": "bit 0: three-waves-per-SIMD kernel, bit 1: doubled persistent grid; trial -1: not tried (that kernel spills registers)"}

    second = None
    if not args.no_second:
        s = Stream(1)
        x2_np = s.points(D, N)
        c2 = 0.1 * s.normal(tri.numCoeffs)
        tri.SetCoeffs(c2)
        x2 = torch.from_numpy(x2_np).cuda()
        log("same map, N(0,I) data, 0.1 N(0,1) coefficients")
        second, _ = measure(tri, x2, args.steps, args.warmup, args.plain_steps, not args.no_plain)
        second["oracle_max_rel_err_2000pts"] = oracle_check(x2_np, c2, gpu_kl)
        second["data"] = "x ~ N(0, I_8) (seed 1 stream), coefficients 0.1 N(0,1) from the same stream"
        if not args.no_input_grad:
            second["kl_input_grad"] = input_grad_workload(tri, x2, max(3, args.plain_steps))
        if not args.no_train:
            second["training"] = training_workload(tri, x2)
        del x2

    many = None
    many_full = None
    if not args.no_many:
        many = many_maps(max(3, args.plain_steps))
        # the full ranges (dims 2..16, orders 2..5, up to 65 536 points); the
        # loop's first call tunes each large map's lanes per point
        many_full = many_maps(2, dmax=16, pmax=5, nmax=65536)
    single = None
    if not args.no_single:
        single = single_component_workload(max(3, args.plain_steps))
    uneven = None
    if not args.no_uneven:
        uneven = uneven_map_workload(2, args.uneven_plain_n)

    kinfo = kernel_info(tri, fused_schedule["choice"])
    th = traffic["headline"] if traffic else None
    if second is not None:
        ta = traffic["active"] if traffic else None
        second["roofline"] = {"bound": "fp64-valu", "achieved": second["achieved_tflops"], "peak": PEAK_FP64_TFLOPS,
                              "unit": "TFLOP/s", "frac": second["achieved_tflops"] / PEAK_FP64_TFLOPS,
                              "regrouped_model_frac": second["regrouped_model_tflops"] / PEAK_FP64_TFLOPS,
                              "traffic": ta["bytes_per_launch"] if ta else None,
                              "algorithmic_bytes_per_launch": ta["algorithmic_bytes"] if ta else None}
    out = {
        "metric": METRIC,
        "value": head["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: banana recursion (seed 0), factory-initial (zero) coefficients",
        "config": {"workload": "KL fwd+adjoint, 8-D triangular map, total order 3 (494 coefficients), "
                               "ProbabilistHermite/SoftPlus/AdaptiveSimpson, N=500000 banana points",
                   "N": N, "dim": D, "order": ORDER, "num_coeffs": tri.numCoeffs},
        "roofline": {"bound": "fp64-valu", "achieved": head["achieved_tflops"], "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": head["achieved_tflops"] / PEAK_FP64_TFLOPS,
                     "traffic": th["bytes_per_launch"] if th else None,
                     "note": "FP64 on the vector ALUs (no MFMA: nothing is a dense contraction); MI355X vector "
                             "FP64 peak 78.6 TFLOP/s; achieved = flops of the reference's algorithm for the "
                             "launch's points and adaptive trees (mpart_amd/roofline.py algo_flops) / HIP-event "
                             "time of the fused kernel + its reduction on the launch stream",
                     "algorithmic_flops_per_launch": head["algorithmic_flops"],
                     "algorithmic_flops_per_point": head["algorithmic_flops_per_point"],
                     "regrouped_model_flops_per_point": head["regrouped_model_flops_per_point"],
                     "regrouped_model_frac": head["regrouped_model_tflops"] / PEAK_FP64_TFLOPS,
                     "algorithmic_bytes_per_launch": th["algorithmic_bytes"] if th else None,
                     "algorithmic_bytes_note": "points 8*D*N + coefficients 8*K + partial rows 8*(K+1) per block",
                     "traffic_detail": traffic, "traffic_error": traffic_err},
        "fused_kernel": kinfo,
        "cpu_baseline": None,
        "fused": head,
        "specialized_kernels": {"active": spec, "compile_s": compile_s, "error": gen_err,
                                "fused_schedule": fused_schedule,
                                "compiled_by_child": prep,
                                "note": "compiled once per run by a child process into MPART_CODEGEN_CACHE; the "
                                        "profiled passes and this process load that code object"},
        "active_quadrature_workload": second,
        "many_maps_workload": many,
        "many_maps_full_ranges_workload": many_full,
        "single_component_workload": single,
        "uneven_map_workload": uneven,
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(x_np, c0, args.cpu_seconds)
    if own_cache:
        import shutil

        shutil.rmtree(own_cache, ignore_errors=True)
    print(json.dumps(out), flush=True)


def cpu_baseline(x_np, coeffs, budget_s):
    """The oracle (CPU restatement of the reference path, OpenMP over points,
    one thread per point as the reference's Kokkos OpenMP backend runs it) on
    the whole headline workload: one untimed run, then the median of five
    timed runs.  Threads: this process's CPU share (OMP_NUM_THREADS when set,
    else the CPUs the process may run on).  budget_s is kept for the command
    line and bounds nothing: six runs take about 10 s on 16 threads."""
    from oracle import oracle as orc

    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS") or share)
    orc.set_threads(threads)
    r = orc.CreateTriangular(D, D, ORDER, orc.options())
    r.SetCoeffs(coeffs)
    xs = np.ascontiguousarray(x_np)
    n = xs.shape[1]
    r.KL(xs)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        r.KL(xs)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n / med, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload ({n} points), KL value + coefficient gradient: median of 5 timed runs "
                      f"after 1 warm-up ({med:.2f} s per run; runs {', '.join(f'{t:.2f}' for t in times)} s)",
            "cpu_model": model, "logical_cpus_on_host": os.cpu_count(), "cpus_available_to_process": share,
            "threads_note": "the GPU box allots 16 host threads per GPU (OMP_NUM_THREADS=16); the oracle runs "
                            "on that share"}


if __name__ == "__main__":
    main()
