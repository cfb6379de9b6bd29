"""Runtime-specialized (hipRTC) kernels against the precompiled generic kernels
and the oracle.

The specialized module is what every map with at most 4096 coefficients uses
by default; SetCodegen(False) selects the generic kernels.  Both compute the
same per-point arithmetic (same basis recurrences, quadrature and acceptance
rule); only the order of the gradient's cross-lane sums differs, so results
agree to 1e-12 relative and the adaptive trees (integrand-evaluation counts)
are identical.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc

from test_gpu_parity import CASES, close, gpu_opts, ref_opts

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _tri(case, dim, order, seed, n, scale=0.1):
    s = Stream(seed)
    g = mt.CreateTriangular(dim, dim, order, gpu_opts(**case))
    x = s.points(dim, n)
    c = scale * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    return g, x, c


def test_specialized_module_is_used():
    _case("codegen_active")
    g = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    spec, secs, err = g.PrepareCodegen()
    print(f"[codegen] C3 compile {secs:.2f} s", flush=True)
    assert spec and err == ""
    # A second map of the same structure reuses the compiled module (an
    # in-process lookup; the first compile may itself have been served from
    # the compiler's on-disk cache of an earlier process, so it can be as fast).
    g2 = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    spec2, secs2, _ = g2.PrepareCodegen()
    assert spec2 and secs2 < max(secs, 0.05)


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_specialized_vs_generic(ci):
    _case("codegen_vs_generic", case=CASES[ci])
    g, x, _ = _tri(CASES[ci], 4, 3, 1000 + ci, n=900)
    sens = Stream(11).points(4, x.shape[1])
    res = {}
    for mode in (True, False):
        g.SetCodegen(mode)
        y, ne = g.EvaluateCounted(x)
        res[mode] = dict(
            y=y, ne=ne, ld=g.LogDeterminant(x), cg=g.CoeffGrad(x, sens), lcg=g.LogDeterminantCoeffGrad(x),
            kl=g.KLValueAndGrad(x), klp=g.KLValueAndGrad(x, plain=True), gr=g.Gradient(x, sens),
            lig=g.LogDeterminantInputGrad(x) if CASES[ci].get("contDeriv", True) else np.zeros(1))
    a, b = res[True], res[False]
    assert np.array_equal(a["ne"], b["ne"])
    for k in ("y", "ld", "cg", "lcg", "gr", "lig"):
        close(a[k], b[k], rtol=1e-12, what=k)
    for k in ("kl", "klp"):
        close(a[k][0], b[k][0], rtol=1e-12, what=k + " loss")
        close(a[k][1], b[k][1], rtol=1e-12, what=k + " grad")


def test_specialized_custom_multiset_vs_oracle():
    # A sparse set with gaps: no constant term, coordinates skipped, mixed orders.
    _case("codegen_custom_set")
    terms = [[(2, 1)], [(0, 2), (2, 1)], [(1, 3)], [(0, 1), (1, 1), (2, 2)], [(2, 3)], [(0, 4)], [(1, 1), (2, 4)]]
    starts, dims, orders = [0], [], []
    for t in terms:
        for d, o in t:
            dims.append(d)
            orders.append(o)
        starts.append(len(dims))
    mset = mt.FixedMultiIndexSet.from_csr(3, starts, dims, orders)
    for case in (CASES[0], CASES[3], CASES[4]):
        g = mt.CreateComponent(mset, gpu_opts(**case))
        r = orc.CreateComponent(ref_opts(**case), 3, starts=starts, dims=dims, orders=orders)
        assert g.PrepareCodegen()[0]
        s = Stream(31)
        x = s.points(3, 600)
        c = 0.3 * s.normal(g.numCoeffs)
        g.SetCoeffs(c)
        r.SetCoeffs(c)
        yg, neg = g.EvaluateCounted(x)
        yr, ner = r.Evaluate(x, return_nevals=True)
        assert np.array_equal(neg, ner)
        close(yg, yr, what="Evaluate")
        close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")
        sens = s.normal(x.shape[1])[None, :]
        close(g.CoeffGrad(x, sens), r.CoeffGrad(x, sens), what="CoeffGrad")
        close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad")


def test_specialized_strided_coefficients():
    torch = pytest.importorskip("torch")
    _case("codegen_strided_coeffs")
    g, x, c = _tri({}, 3, 3, 1100, n=500)
    ref_y = g.Evaluate(x)
    ref_l, ref_g = g.KLValueAndGrad(x)
    buf = torch.zeros(2 * g.numCoeffs, dtype=torch.float64, device="cuda")
    buf[::2] = torch.from_numpy(c).cuda()
    g.WrapCoeffs(buf[::2])
    assert np.array_equal(g.Evaluate(x), ref_y)
    l2, g2 = g.KLValueAndGrad(x)
    close(l2, ref_l, rtol=1e-14, what="loss")
    close(g2, ref_g, rtol=1e-14, what="grad")


def test_large_map_falls_back_to_generic():
    _case("codegen_fallback")
    g = mt.CreateTotalOrderComponent(8, 7, mt.MapOptions())  # 6435 coefficients
    spec, _, err = g.PrepareCodegen()
    assert not spec and "4096" in err
    r = orc.CreateComponent(ref_opts(), 8, 7)
    s = Stream(41)
    x = s.points(8, 64)
    c = 0.01 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    close(g.Evaluate(x), r.Evaluate(x), what="Evaluate")


def test_fused_schedule_autotuning():
    torch = pytest.importorskip("torch")
    _case("codegen_autotune")
    g = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    s = Stream(17)
    x = torch.from_numpy(s.points(8, 70000)).cuda()
    g.SetCoeffs(0.1 * s.normal(g.numCoeffs))
    assert g.FusedSchedule()[0] == -1
    l1, g1 = g.KLValueAndGrad(x)
    choice, trials = g.FusedSchedule()
    print(f"[autotune] choice {choice}, trial ms {['%.3f' % t for t in trials]}", flush=True)
    # schedules whose kernel would spill registers are not tried (-1)
    assert 0 <= choice <= 3 and trials[choice] > 0
    assert trials[0] > 0 and trials[2] > 0 and all(t > 0 or t == -1.0 for t in trials)
    lp, gp = g.KLValueAndGrad(x, plain=True)
    close(l1.item(), lp.item(), rtol=1e-12, what="loss")
    close(g1.cpu().numpy(), gp.cpu().numpy(), rtol=1e-12, what="grad")
    l2, g2 = g.KLValueAndGrad(x)  # the tuned schedule, deterministic from here on
    l3, g3 = g.KLValueAndGrad(x)
    assert torch.equal(g2, g3) and torch.equal(l2, l3)


def test_fused_kl_step_replays_in_a_graph():
    # One training step's objective + gradient captured once and replayed with
    # new coefficients written into the wrapped (static) coefficient tensor.
    torch = pytest.importorskip("torch")
    _case("codegen_graph_step")
    g = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    s = Stream(23)
    x = torch.from_numpy(s.points(8, 100000)).cuda()
    coeffs = torch.from_numpy(0.05 * s.normal(g.numCoeffs)).cuda()
    g.WrapCoeffs(coeffs)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g.KLValueAndGradImpl(x, loss, grad)  # compiles, tunes and sizes the work buffers
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        g.KLValueAndGradImpl(x, loss, grad)
    for step in range(3):
        new = torch.from_numpy((0.05 + 0.01 * step) * Stream(30 + step).normal(g.numCoeffs)).cuda()
        coeffs.copy_(new)
        loss.fill_(-1.0)
        grad.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        le = torch.zeros_like(loss)
        ge = torch.zeros_like(grad)
        g.KLValueAndGradImpl(x, le, ge)  # eager, same buffers and schedule
        torch.cuda.synchronize()
        assert torch.equal(loss, le) and torch.equal(grad, ge), f"step {step}"



def test_forced_fused_schedules():
    # Each schedule sums the points in its own order: 1e-13 relative to each
    # other and 1e-10 to the oracle; a forced schedule is kept.
    torch = pytest.importorskip("torch")
    _case("codegen_forced_schedules")
    g = mt.CreateTriangular(6, 6, 3, mt.MapOptions())
    r = orc.CreateTriangular(6, 6, 3, orc.options())
    s = Stream(41)
    xn = s.points(6, 6000)
    x = torch.from_numpy(xn).cuda()
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    out = {}
    for ch in (0, 1, 2, 3):
        g.SetFusedSchedule(ch)
        l, gr = g.KLValueAndGrad(x)
        out[ch] = (l.cpu().numpy().reshape(1), gr.cpu().numpy())
        assert g.FusedSchedule()[0] == ch
    for ch in (1, 2, 3):
        close(out[ch][0], out[0][0], rtol=1e-13, what=f"loss, schedule {ch}")
        close(out[ch][1], out[0][1], rtol=1e-13, what=f"gradient, schedule {ch}")
    lr, gr_ = r.KL(xn)
    close(out[2][0], np.array([lr]), what="loss vs oracle")
    close(out[2][1], gr_, what="gradient vs oracle")
    with pytest.raises(ValueError):
        g.SetFusedSchedule(4)
    g.SetFusedSchedule(-1)
    assert g.FusedSchedule()[0] == -1
