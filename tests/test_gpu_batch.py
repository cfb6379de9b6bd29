"""Batched execution of many independent maps (one persistent launch per
method) and its capture in a HIP graph.

Maps: triangular maps with dimension and total order drawn from a seeded
stream (a scaled-down version of the many-maps workload: dims 2..9, orders
2..4, 100..700 points each) plus single components.  The batched kernel runs
the same per-point arithmetic as the single-map generic kernels, so its
results equal theirs bitwise; they are also checked against the oracle.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc

from test_gpu_parity import close

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu-case] {name}", flush=True)


def _make_batch(nmaps=20, seed=11):
    s = Stream(seed)
    u = s.uniform(4 * nmaps).reshape(nmaps, 4)
    maps, refs, xs, tri = [], [], [], []
    for i in range(nmaps):
        d = 2 + int(u[i, 0] * 8)
        p = 2 + int(u[i, 1] * 3)
        n = 100 + int(u[i, 2] * 600)
        if i % 5 == 4:  # a single component among the triangular maps
            g = mt.CreateTotalOrderComponent(d, p, mt.MapOptions())
            r = orc.CreateComponent(orc.options(), d, p)
        else:
            g = mt.CreateTriangular(d, d, p, mt.MapOptions())
            r = orc.CreateTriangular(d, d, p, orc.options())
        c = 0.1 * s.normal(g.numCoeffs)
        g.SetCoeffs(c)
        r.SetCoeffs(c)
        maps.append(g)
        refs.append(r)
        xs.append(torch.from_numpy(s.points(d, n)).cuda())
        tri.append(i % 5 != 4)
    return maps, refs, xs, tri


@pytest.fixture(scope="module")
def batch():
    maps, refs, xs, tri = _make_batch()
    return mt.MapBatch(maps), maps, refs, xs, tri


def _as_dev(a):
    return a if isinstance(a, torch.Tensor) else torch.from_numpy(a).cuda()


def test_batched_evaluate_and_logdet(batch):
    _case("batch_evaluate")
    b, maps, refs, xs, _ = batch
    ys, lds = b.Evaluate(xs, logdet=True)
    for m, r, x, y, ld in zip(maps, refs, xs, ys, lds):
        m.SetCodegen(False)  # the single-map generic kernels: same arithmetic, same bits
        assert torch.equal(y, _as_dev(m.Evaluate(x)))
        assert torch.equal(ld, _as_dev(m.LogDeterminant(x)))
        m.SetCodegen(True)
        xn = x.cpu().numpy()
        close(y.cpu().numpy(), r.Evaluate(xn), what="batched Evaluate")
        close(ld.cpu().numpy(), r.LogDeterminant(xn), what="batched LogDeterminant")


def _sens(maps, xs, seed):
    s = Stream(seed)
    return [torch.from_numpy(s.normal(m.outputDim * x.shape[1]).reshape(m.outputDim, x.shape[1])).cuda()
            for m, x in zip(maps, xs)]


def test_batched_coeff_grad(batch):
    _case("batch_coeff_grad")
    b, maps, refs, xs, _ = batch
    sens = _sens(maps, xs, 12)
    outs = b.CoeffGrad(xs, sens)
    for m, r, x, se, o in zip(maps, refs, xs, sens, outs):
        m.SetCodegen(False)
        single = _as_dev(m.CoeffGrad(x, se))
        m.SetCodegen(True)
        assert torch.equal(o, single)
        close(o.cpu().numpy(), r.CoeffGrad(x.cpu().numpy(), se.cpu().numpy()), what="batched CoeffGrad")


def _inverse_args(maps, refs, xs, tri):
    x1s, rs = [], []
    for m, r, x, t in zip(maps, refs, xs, tri):
        rs.append(torch.from_numpy(r.Evaluate(x.cpu().numpy())).cuda())
        # triangular map: no conditioning rows; component: all rows, the last one the initial guess
        x1s.append(x[:0] if t else x)
    return x1s, rs


def test_batched_inverse(batch):
    # Schedule 1 (per-lane solve scheduling) runs the single-map path's setup
    # and solver, so its roots equal the single-map Inverse's bit for bit
    # (maps with many terms take 4 lanes per point in the setup of both).
    _case("batch_inverse")
    b, maps, refs, xs, tri = batch
    x1s, rs = _inverse_args(maps, refs, xs, tri)
    b.SetInverseSchedule(1)
    outs = b.Inverse(x1s, rs)
    b.SetInverseSchedule(0)
    for m, r, x, x1, rr, o, t in zip(maps, refs, xs, x1s, rs, outs, tri):
        single = _as_dev(m.Inverse(x1, rr))
        assert torch.equal(o, single)
        ref = r.Inverse(x1.cpu().numpy(), rr.cpu().numpy())
        got = o.cpu().numpy()
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(got))
        if t:
            # initial guesses 0: agreement to the oracle's arithmetic
            assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-9
        else:
            # initial guess = the solution: agreement to the solver tolerance
            assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-5


def test_batched_sequence_replays_in_one_graph(batch):
    # Evaluate + LogDeterminant, CoeffGrad and Inverse captured in one graph and
    # replayed three times with eager work in between; every replay must give
    # the eager results.
    _case("batch_graph")
    b, maps, refs, xs, tri = batch
    ys = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    lds = [torch.empty(x.shape[1], dtype=torch.float64, device="cuda") for x in xs]
    sens = _sens(maps, xs, 13)
    cgs = [torch.empty((m.numCoeffs, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    x1s, rs = _inverse_args(maps, refs, xs, tri)
    xis = [torch.empty_like(r) for r in rs]

    def seq():
        b.EvaluateImpl(xs, ys, lds)
        b.CoeffGradImpl(xs, sens, cgs)
        b.InverseImpl(x1s, rs, xis)

    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        seq()  # eager: uploads each method's arguments
    torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    ref = [[t.clone() for t in ts] for ts in (ys, lds, cgs, xis)]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        seq()
    for rep in range(3):
        for ts in (ys, lds, cgs, xis):
            for t in ts:
                t.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for ts, es in zip((ys, lds, cgs, xis), ref):
            for a, e in zip(ts, es):
                assert torch.equal(a, e), f"replay {rep}"
        # eager work between replays: unrelated torch work, and an eager call
        # with the captured arguments (a call with other arguments would
        # redirect the graph, which reads the method's arguments at replay)
        scratch = torch.randn(1 << 20, dtype=torch.float64, device="cuda")
        (scratch * 2.0).sum().item()
        b.EvaluateImpl(xs, ys, lds)
        torch.cuda.synchronize()
        for a, e in zip(ys, ref[0]):
            assert torch.equal(a, e)


def test_batch_rejects_mixed_options():
    _case("batch_mixed")
    a = mt.CreateTriangular(2, 2, 2, mt.MapOptions())
    c = mt.CreateTriangular(2, 2, 2, mt.MapOptions(posFuncType=mt.PosFuncTypes.Exp))
    with pytest.raises(NotImplementedError):
        mt.MapBatch([a, c])


def test_batched_evaluate_group_schedule():
    # The mixed schedule: maps with many terms take 4 lanes per point (a
    # shared basis table per point, the terms split over the lanes), the
    # others a lane per point, in one launch.  Same values as the per-map
    # kernels up to the order of the C_q sums, checked against the oracle.
    _case("batch_eval_group")
    maps, refs, xs, tri = _make_batch(nmaps=12, seed=13)
    s = Stream(14)
    for d, p in ((11, 4), (9, 5)):  # 4367 and 5004 coefficients: 4 lanes per point
        g = mt.CreateTriangular(d, d, p, mt.MapOptions())
        r = orc.CreateTriangular(d, d, p, orc.options())
        c = 0.05 * s.normal(g.numCoeffs)
        g.SetCoeffs(c)
        r.SetCoeffs(c)
        maps.append(g)
        refs.append(r)
        xs.append(torch.from_numpy(s.points(d, 333)).cuda())
    b = mt.MapBatch(maps)
    b.SetPointGroup(4)
    ys = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    lds = [torch.empty(x.shape[1], dtype=torch.float64, device="cuda") for x in xs]
    b.EvaluateImpl(xs, ys, lds)
    torch.cuda.synchronize()
    for m, r, x, y, ld in zip(maps, refs, xs, ys, lds):
        xn = x.cpu().numpy()
        close(y.cpu().numpy(), r.Evaluate(xn), what="batched Evaluate, mixed schedule")
        close(ld.cpu().numpy(), r.LogDeterminant(xn), what="batched LogDeterminant, mixed schedule")
        close(y.cpu().numpy(), m.Evaluate(x).cpu().numpy(), rtol=1e-12, what="group vs per-map")
    assert b.EvalSchedule()[0] == 4


def test_batched_evaluate_tuner_and_capture_freeze():
    # Enough points to tune (>= 65536 in all); after a captured call the
    # schedule cannot change.
    _case("batch_eval_tuner")
    s = Stream(17)
    maps, xs = [], []
    for i in range(6):
        d, p = 4 + 2 * i, 3 if i % 2 else 4
        m = mt.CreateTriangular(d, d, p, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        m.SetCodegen(False)
        maps.append(m)
        xs.append(torch.from_numpy(s.points(d, 12000)).cuda())
    b = mt.MapBatch(maps)
    ys = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    lds = [torch.empty(x.shape[1], dtype=torch.float64, device="cuda") for x in xs]
    b.EvaluateImpl(xs, ys, lds)
    torch.cuda.synchronize()
    lanes, trials = b.EvalSchedule()
    print(f"[batch] tuned lanes {lanes} trials {trials}", flush=True)
    assert lanes in (1, 4) and all(t > 0 for t in trials)
    for m, x, y, ld in zip(maps, xs, ys, lds):
        close(y.cpu().numpy(), m.Evaluate(x).cpu().numpy(), rtol=1e-12, what="tuned batch vs per-map")
        close(ld.cpu().numpy(), m.LogDeterminant(x).cpu().numpy(), rtol=1e-12, what="tuned logdet vs per-map")
    y0 = [t.clone() for t in ys]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.EvaluateImpl(xs, ys, lds)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.EvaluateImpl(xs, ys, lds)
    for _ in range(3):
        for t in ys:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(a, c) for a, c in zip(ys, y0))
    with pytest.raises(ValueError):
        b.SetPointGroup(5 - lanes)  # the other schedule


def test_batched_inverse_schedule_tuner():
    # >= 65536 points in all: the first call times both inverse schedules and
    # keeps the faster; its roots equal those of a batch forced to that schedule.
    _case("batch_inverse_tuner")
    s = Stream(31)
    maps, xs, ys = [], [], []
    for d, p in ((3, 2), (4, 3), (5, 2), (2, 4)):
        m = mt.CreateTriangular(d, d, p, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        x = torch.from_numpy(s.points(d, 20_000)).cuda()
        maps.append(m)
        xs.append(x)
        ys.append(_as_dev(m.Evaluate(x)))
    x1s = [x[:0] for x in xs]
    b = mt.MapBatch(maps)
    assert b.InverseSchedule()[0] == 0
    got = b.Inverse(x1s, ys)
    sched, trial = b.InverseSchedule()
    assert sched in (1, 2) and min(trial) > 0.0
    forced = mt.MapBatch(maps)
    forced.SetInverseSchedule(sched)
    ref = forced.Inverse(x1s, ys)
    for a, c in zip(got, ref):
        assert torch.equal(torch.isnan(a), torch.isnan(c))
        assert torch.equal(torch.nan_to_num(a), torch.nan_to_num(c))
    # both schedules solve to the same tolerance: T(x) is recovered where T is well conditioned
    other = mt.MapBatch(maps)
    other.SetInverseSchedule(3 - sched)
    for a, c in zip(got, other.Inverse(x1s, ys)):
        fin = torch.isfinite(a) & torch.isfinite(c)
        assert torch.equal(torch.isfinite(a), torch.isfinite(c))
        assert (a[fin] - c[fin]).abs().max().item() <= 1e-9


def test_captured_queue_inverse_rejects_other_arguments():
    # The per-lane Inverse schedule bakes its plan's offsets into the captured
    # launches: once captured, a call with other arguments is refused on the
    # host (before any launch), and the graph still replays to the eager roots.
    _case("batch_inverse_capture_guard")
    s = Stream(41)
    maps, xs = [], []
    for d, p in ((3, 2), (4, 2)):
        m = mt.CreateTriangular(d, d, p, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        maps.append(m)
        xs.append(torch.from_numpy(s.points(d, 3000)).cuda())
    b = mt.MapBatch(maps)
    b.SetInverseSchedule(1)
    ys = b.Evaluate(xs)
    x1s = [x[:0] for x in xs]
    outs = [torch.empty_like(x) for x in xs]
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        b.InverseImpl(x1s, ys, outs)
    torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.InverseImpl(x1s, ys, outs)
    other = [torch.empty_like(x) for x in xs]
    with pytest.raises(ValueError, match="captured"):
        b.InverseImpl(x1s, ys, other)
    with pytest.raises(ValueError, match="captured"):
        b.SetInverseSchedule(2)
    for o in outs:
        o.fill_(-1.0)
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for o, e in zip(outs, ref):
        assert torch.equal(torch.nan_to_num(o), torch.nan_to_num(e))


def test_queue_inverse_mixed_order_classes():
    # Orders 2..5 in one batch: the batch runs every map with the kernels of the
    # largest last-coordinate order class, the per-map path with each map's own;
    # a single component with a one-column x1 rides along.  Forced per-lane
    # schedule against the per-map roots and the oracle (1e-9).
    _case("batch_inverse_queue_mixed_orders")
    s = Stream(51)
    maps, refs, xs, x1s, ys, tri = [], [], [], [], [], []
    for d, p, t in ((2, 5, True), (3, 2, True), (4, 3, True), (3, 4, False), (2, 2, True)):
        if t:
            g = mt.CreateTriangular(d, d, p, mt.MapOptions())
            r = orc.CreateTriangular(d, d, p, orc.options())
        else:
            g = mt.CreateTotalOrderComponent(d, p, mt.MapOptions())
            r = orc.CreateComponent(orc.options(), d, p)
        c = 0.1 * s.normal(g.numCoeffs)
        g.SetCoeffs(c)
        r.SetCoeffs(c)
        n = 700 + 97 * len(maps)
        x = torch.from_numpy(s.points(d, n)).cuda()
        maps.append(g)
        refs.append(r)
        xs.append(x)
        ys.append(torch.from_numpy(r.Evaluate(x.cpu().numpy())).cuda())
        x1s.append(x[:0] if t else x[:, :1].contiguous())
        tri.append(t)
    b = mt.MapBatch(maps)
    b.SetInverseSchedule(1)
    outs = b.Inverse(x1s, ys)
    for m, r, x1, y, o, t in zip(maps, refs, x1s, ys, outs, tri):
        single = _as_dev(m.Inverse(x1, y)).cpu().numpy()
        ref = r.Inverse(x1.cpu().numpy(), y.cpu().numpy())
        got = o.cpu().numpy()
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(got))
        assert np.array_equal(np.isfinite(single), np.isfinite(got))
        assert np.max(np.abs(got[fin] - ref[fin]), initial=0.0) <= 1e-9
        assert np.max(np.abs(got[fin] - single[fin]), initial=0.0) <= 1e-9


def _graph_of(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()  # eager first: tunes, sizes and uploads
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


def test_batch_arguments_frozen_after_capture():
    # A captured batched launch holds its tables, tile count and grid: a later
    # call of that method with other (here larger) arguments is refused before
    # anything is launched, and the graph keeps replaying the captured call.
    _case("batch_capture_frozen")
    maps, _, xs, _ = _make_batch(nmaps=6, seed=41)
    b = mt.MapBatch(maps)
    ys = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, xs)]
    g = _graph_of(lambda: b.EvaluateImpl(xs, ys))
    torch.cuda.synchronize()
    ref = [y.clone() for y in ys]
    s = Stream(42)
    big = [torch.from_numpy(s.points(x.shape[0], 4 * x.shape[1] + 1000)).cuda() for x in xs]
    ybig = [torch.empty((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, big)]
    with pytest.raises(ValueError, match="captured"):
        b.EvaluateImpl(big, ybig)
    for y in ys:
        y.fill_(-1.0)
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for a, e in zip(ys, ref):
        assert torch.equal(a, e)
    # the captured arguments are still accepted eagerly
    b.EvaluateImpl(xs, ys)
    torch.cuda.synchronize()
    for a, e in zip(ys, ref):
        assert torch.equal(a, e)
    # another method of the same batch is not frozen until it is captured
    cgs = [torch.empty((m.numCoeffs, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, big)]
    sens = [torch.ones((m.outputDim, x.shape[1]), dtype=torch.float64, device="cuda") for m, x in zip(maps, big)]
    b.CoeffGradImpl(big, sens, cgs)
    torch.cuda.synchronize()


def test_map_buffers_outlive_capture():
    # A captured single-map Inverse holds the map's work buffers (per-point
    # records, the triangular point buffer).  A later eager call with more
    # points needs larger buffers: the map allocates new ones and keeps the
    # captured ones, so the graph still replays to the captured results.
    _case("map_capture_buffers")
    m = mt.CreateTriangular(3, 3, 2, mt.MapOptions())
    s = Stream(43)
    m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
    x = torch.from_numpy(s.points(3, 3000)).cuda()
    y = m.Evaluate(x)
    x1 = x[:0]
    xi = torch.empty_like(x)
    g = _graph_of(lambda: m.InverseImpl(x1, y, xi))
    torch.cuda.synchronize()
    ref = xi.clone()
    xb = torch.from_numpy(s.points(3, 60000)).cuda()
    yb = m.Evaluate(xb)
    xib = m.Inverse(xb[:0], yb)
    torch.cuda.synchronize()
    assert torch.isfinite(xib).any()
    xi.fill_(-1.0)
    torch.cuda.synchronize()
    for _ in range(2):
        g.replay()
        torch.cuda.synchronize()
        assert bool(((xi == ref) | (torch.isnan(xi) & torch.isnan(ref))).all())
