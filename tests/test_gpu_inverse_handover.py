"""Hand-over of long inverse solves to the one-wavefront-per-solve pass.

A queue lane hands its solve over once its wavefront finds no point left to
take; with the budget at one step, that is every solve still running then
(up to the list's capacity; with a few hundred points, nearly all of them),
and each is solved again by a wavefront whose 64 lanes build each adaptive
quadrature tree level by level.  The trees, node values and sums are
those of the one-lane depth-first loop, so the roots must equal the roots
with the hand-over off bit for bit (NaN where those are NaN), and the
oracle's to 1e-9.  Cases: option sets with adaptive Simpson (plain, Hermite
functions with Exp, nugget without normalization, a forced acceptance level,
LinearizedBasis bounds) and adaptive Clenshaw-Curtis (5- and 9-point rules,
PhysicistHermite, LinearizedBasis bounds, a 33-point rule in device
memory), NaN inputs, targets outside the range (NaN roots
after a long bracket search), triangular maps (one pass per component), the
batched queue (schedule 1), an overflowing integrand (the pass falls back to
the depth-first loop, which records the status word), trees larger than the
wavefront's workspace (the same fallback), and the default budget on a
large input.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc

from test_gpu_parity import CASES, gpu_opts, ref_opts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

# option sets with an adaptive rule: Simpson (CASES 0, 2, 6, 7, 12) and nested
# Clenshaw-Curtis (CASES 4, 5, 9, 13, and a rule of 33 fine points read from
# device memory)
ADAPTIVE = [CASES[i] for i in (0, 2, 6, 7, 12, 4, 5, 9, 13)] + [
    dict(quadType="AdaptiveClenshawCurtis", quadPts=20, quadAbsTol=1e-9, quadRelTol=1e-9)]


def _case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _same(a, b):
    assert torch.equal(torch.isnan(a), torch.isnan(b))
    assert torch.equal(torch.nan_to_num(a), torch.nan_to_num(b))


def _component(case, seed, n, dim=3, order=3, scale=0.2):
    s = Stream(seed)
    g = mt.CreateTotalOrderComponent(dim, order, gpu_opts(**case))
    x = s.points(dim, n)
    c = scale * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    tgt = 3.0 * s.normal(n)[None, :]
    x[0, 3] = np.nan
    return g, c, x, tgt


def _close_roots(got, ref):
    """Roots against the oracle's: 1e-9, except where a comparison of T with
    the target is decided by the last bits of T (the device and the host
    round a few operations differently): there the two solvers take
    different, equally valid ITP paths and stop at different points of the
    tolerance window (xtol = 1e-6).  At most 1 % of the points, each within
    2 xtol."""
    d = np.abs(got - ref)
    off = int(np.count_nonzero(d > 1e-9))
    print(f"[gpu-case] roots off the oracle's by more than 1e-9: {off} of {d.size}", flush=True)
    assert off <= max(1, d.size // 100)
    assert np.max(d) <= 2e-6


@pytest.mark.parametrize("ci", range(len(ADAPTIVE)))
@pytest.mark.parametrize("budget", [1, 64])
def test_handed_over_solves_equal_one_lane_solves(ci, budget):
    case = ADAPTIVE[ci]
    _case("handover", case=case, budget=budget)
    g, c, x, tgt = _component(case, 1800 + ci, 900)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    g.SetInverseHandover(0)
    off = g.Inverse(xd, td)
    g.SetInverseHandover(budget)
    on = g.Inverse(xd, td)
    _same(on, off)
    st = g.InverseHandoverStats()
    print(f"[gpu-case] handover stats {st}", flush=True)
    assert st["solves"] > (100 if budget == 1 else 0)
    r = orc.CreateComponent(ref_opts(**case), 3, 3)
    r.SetCoeffs(c)
    ref = r.Inverse(x, tgt)
    got = on.cpu().numpy()
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.isnan(got[0, 3])
    _close_roots(got[fin], ref[fin])


def test_handover_list_full_and_triangular():
    # 6 000 solves against a list of 1 024: the rest finish in the queue.
    _case("handover_list_full")
    g, c, x, tgt = _component({}, 1850, 6000)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    g.SetInverseHandover(0)
    off = g.Inverse(xd, td)
    g.SetInverseHandover(1)
    _same(g.Inverse(xd, td), off)
    # a triangular map: one queue pass and one hand-over pass per component
    _case("handover_triangular")
    s = Stream(1851)
    t = mt.CreateTriangular(3, 3, 3, mt.MapOptions())
    t.SetCoeffs(0.1 * s.normal(t.numCoeffs))
    n = 3000
    y = torch.from_numpy(2.0 * s.normal(3 * n).reshape(3, n)).cuda()
    empty = torch.zeros((0, n), dtype=torch.float64, device="cuda")
    t.SetInverseHandover(0)
    toff = t.Inverse(empty, y)
    t.SetInverseHandover(1)
    _same(t.Inverse(empty, y), toff)


def test_handover_batched_queue():
    _case("handover_batch")
    s = Stream(1860)
    maps, ys = [], []
    for dim, order, n in [(2, 2, 700), (3, 3, 1200), (4, 2, 500)]:
        m = mt.CreateTriangular(dim, dim, order, mt.MapOptions())
        m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
        maps.append(m)
        ys.append(torch.from_numpy(2.0 * s.normal(dim * n).reshape(dim, n)).cuda())
    x1 = [torch.zeros((0, y.shape[1]), dtype=torch.float64, device="cuda") for y in ys]
    b = mt.MapBatch(maps)
    b.SetInverseSchedule(1)
    b.SetInverseHandover(0)
    off = [o.clone() for o in b.Inverse(x1, ys)]
    b.SetInverseHandover(1)
    on = b.Inverse(x1, ys)
    for a, c in zip(on, off):
        _same(a, c)
    # and the one-lane solver of each map (schedule 2)
    b.SetInverseSchedule(2)
    seq = b.Inverse(x1, ys)
    for a, c in zip(on, seq):
        _same(a, c)


def test_handover_overflowing_integrand_records_status():
    # Exp bijector, coefficient 500 on He_2: g overflows for y > 1.59, so the
    # bracket search for the target 1e200 meets an infinite g; the wavefront
    # pass falls back to the depth-first loop there, and the call raises as
    # with the hand-over off.
    _case("handover_overflow")
    o = mt.MapOptions(posFuncType=mt.PosFuncTypes.Exp)
    g = mt.CreateTotalOrderComponent(1, 2, o)
    g.SetCoeffs(np.array([0.0, 0.0, 500.0]))
    x1 = np.zeros((1, 1))
    for budget in [0, 1]:
        g.SetInverseHandover(budget)
        small = g.Inverse(x1, np.array([[1e30]]))
        with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
            g.Inverse(x1, np.array([[1e200]]))
        g.CheckIntegrand()
        if budget == 0:
            ref_small = small
        else:
            assert np.array_equal(small, ref_small)


def test_handover_clenshaw_curtis_large():
    # The single-component workload's options (dim 6, order 4, adaptive
    # Clenshaw-Curtis, tolerances 1e-8) with N(0, 1) targets: about 1 % of the
    # solves end in NaN after long bracket searches, the ones handed over.
    _case("handover_cc_large")
    case = dict(quadType="AdaptiveClenshawCurtis", quadAbsTol=1e-8, quadRelTol=1e-8)
    s = Stream(1880)
    g = mt.CreateTotalOrderComponent(6, 4, gpu_opts(**case))
    g.SetCoeffs(0.05 * s.normal(g.numCoeffs))
    n = 100_000
    xd = torch.from_numpy(s.points(6, n)).cuda()
    td = torch.from_numpy(s.normal(n)[None, :]).cuda()
    g.SetInverseHandover(0)
    off = g.Inverse(xd, td)
    g.SetInverseHandover()
    _same(g.Inverse(xd, td), off)
    st = g.InverseHandoverStats()
    print(f"[gpu-case] handover stats {st}", flush=True)
    assert st["solves"] > 0 and st["fallback_nonfinite"] == 0


def test_default_handover_large_input():
    # The default budget on 200 000 wide targets (dims 4, order 3): the
    # queue hands over its longest solves; roots equal the one-lane solver's.
    _case("handover_default_large")
    g, c, x, tgt = _component({}, 1870, 200_000, dim=4, order=3, scale=0.15)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    g.SetInverseHandover(0)
    off = g.Inverse(xd, td)
    g.SetInverseHandover()
    _same(g.Inverse(xd, td), off)
    with pytest.raises(ValueError):
        g.SetInverseHandover(-1)


def test_handover_trees_beyond_the_workspace():
    # Tolerances of 1e-13 make the quadrature trees larger than the
    # wavefront's LDS workspace (320 intervals): those T evaluations fall back
    # to the depth-first loop in every lane, with the same roots bit for bit.
    case = dict(quadAbsTol=1e-13, quadRelTol=1e-13, quadMaxSub=24)
    _case("handover_workspace_fallback", case=case)
    g, c, x, tgt = _component(case, 1890, 300, scale=0.3)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    g.SetInverseHandover(0)
    off = g.Inverse(xd, td)
    g.SetInverseHandover(1)
    on = g.Inverse(xd, td)
    _same(on, off)
    st = g.InverseHandoverStats()
    print(f"[gpu-case] handover stats {st}", flush=True)
    assert st["solves"] > 0 and st["fallback_size"] > 0
