"""Group-per-point schedules of the generic Evaluate / LogDeterminant kernels
(4, 8 or 64 lanes per point sharing one basis table) against the oracle and the
lane-per-point kernel, and the autotuner that picks among them.

The group kernels sum each C_q over the lanes of a group, so values differ
from the lane-per-point kernel by rounding only (bound 1e-12 relative); the
adaptive trees (integrand-evaluation counts) are the oracle's.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream, student_t_points
from oracle import oracle as orc

from test_gpu_parity import CASES, close, gpu_opts, ref_opts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("ci", [0, 2, 4, 5, 7, 8, 10, 12, 13])
@pytest.mark.parametrize("lanes", [4, 8, 64])
def test_group_schedule_vs_oracle(ci, lanes):
    _case("point_group", case=CASES[ci], lanes=lanes)
    s = Stream(2000 + ci)
    g = mt.CreateTriangular(5, 5, 3, gpu_opts(**CASES[ci]))
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    g.SetCodegen(False)
    g.SetPointGroup(lanes)
    x = s.points(5, 700)
    r = orc.CreateTriangular(5, 5, 3, ref_opts(**CASES[ci]))
    r.SetCoeffs(c)
    y, ne = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    assert np.array_equal(ne, ner)
    close(y, yr, what="Evaluate")
    if CASES[ci].get("contDeriv", True):
        close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")
    # the lane-per-point kernel on the same points
    g.SetPointGroup(1)
    close(y, g.Evaluate(x), rtol=1e-12, what="group vs lane per point")


def test_uneven_32d_map_all_schedules_and_tuner():
    # Component k of total order max(1, 6 - k // 6) (62 572 coefficients),
    # HermiteFunctions / Exp / adaptive Simpson, Student-t(5) points.
    _case("point_group_32d")
    orders = [max(1, 6 - k // 6) for k in range(32)]
    opts = dict(basisType="HermiteFunctions", posFuncType="Exp")
    g = mt.TriangularMap([mt.CreateTotalOrderComponent(k + 1, orders[k], gpu_opts(**opts)) for k in range(32)])
    c = 0.01 * Stream(5).normal(g.numCoeffs)
    g.SetCoeffs(c)
    r = orc.CreateTriangularOrders(ref_opts(**opts), list(range(1, 33)), orders)
    r.SetCoeffs(c)
    x = student_t_points(32, 200, nu=5, seed=7)
    yr, ner = r.Evaluate(x, return_nevals=True)
    ldr = r.LogDeterminant(x)
    for lanes in (1, 2, 4, 8, 64):
        g.SetPointGroup(lanes)
        y, ne = g.EvaluateCounted(x)
        assert np.array_equal(ne, ner), lanes
        close(y, yr, what=f"Evaluate lanes={lanes}")
        close(g.LogDeterminant(x), ldr, what=f"LogDeterminant lanes={lanes}")
    # the tuner: enough points to tune, and the tuned result matches lane per point
    g.SetPointGroup(0)
    xl = torch.from_numpy(student_t_points(32, 20000, nu=5, seed=8)).cuda()
    y = g.Evaluate(xl)
    ld = g.LogDeterminant(xl)
    (le, ll, _), trials = g.EvalSchedule()
    print(f"[point_group] tuned lanes evaluate={le} logdet={ll} trials={trials}", flush=True)
    assert le in (1, 2, 4, 8, 64) and ll in (1, 2, 4, 8, 64)
    assert all(t > 0 for t in trials[0] + trials[1])
    g.SetPointGroup(1)
    close(y.cpu().numpy(), g.Evaluate(xl).cpu().numpy(), rtol=1e-12, what="tuned vs lane per point")
    close(ld.cpu().numpy(), g.LogDeterminant(xl).cpu().numpy(), rtol=1e-12, what="tuned logdet vs lane per point")


def test_point_group_argument_checked():
    g = mt.CreateTriangular(2, 2, 2, mt.MapOptions())
    with pytest.raises(ValueError):
        g.SetPointGroup(3)


@pytest.mark.parametrize("ci", [0, 1, 2, 6, 9, 13])
@pytest.mark.parametrize("lanes", [2, 4, 8, 64])
def test_group_fused_vs_oracle(ci, lanes):
    # KL value + gradient and the coefficient VJP of the generic fused pass
    # with 4 / 8 / 64 lanes per point, or (2) the components split over the
    # wavefronts of a block, against the oracle and the lane-per-point pass;
    # 701 points leave part of the last wavefront step (tile) without a point.
    _case("point_group_fused", case=CASES[ci], lanes=lanes)
    s = Stream(2100 + ci)
    g = mt.CreateTriangular(4, 4, 3, gpu_opts(**CASES[ci]))
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    g.SetCodegen(False)
    x = np.ascontiguousarray(s.points(4, 701))
    r = orc.CreateTriangular(4, 4, 3, ref_opts(**CASES[ci]))
    r.SetCoeffs(c)
    lr, gr = r.KL(x)
    g.SetPointGroup(lanes)
    lf, gf = g.KLValueAndGrad(x)
    close(lf, lr, what="KL loss")
    close(gf, gr, what="KL grad")
    xd = torch.from_numpy(x).cuda()
    sens = torch.from_numpy(s.normal(4 * 701).reshape(4, 701)).cuda()
    ls = torch.from_numpy(s.normal(701)).cuda()
    v = g.CoeffVJP(xd, sens, ls).cpu().numpy()
    g.SetPointGroup(1)
    close(v, g.CoeffVJP(xd, sens, ls).cpu().numpy(), rtol=1e-12, what="VJP group vs lane per point")
    # a handful of points (fewer than one wavefront step)
    g.SetPointGroup(lanes)
    lf5, gf5 = g.KLValueAndGrad(x[:, :5])
    lr5, gr5 = r.KL(np.ascontiguousarray(x[:, :5]))
    close(lf5, lr5, what="KL loss 5 points")
    close(gf5, gr5, what="KL grad 5 points")


def test_uneven_32d_map_fused_schedules_and_tuner():
    _case("point_group_32d_fused")
    orders = [max(1, 6 - k // 6) for k in range(32)]
    opts = dict(basisType="HermiteFunctions", posFuncType="Exp")
    g = mt.TriangularMap([mt.CreateTotalOrderComponent(k + 1, orders[k], gpu_opts(**opts)) for k in range(32)])
    c = 0.01 * Stream(5).normal(g.numCoeffs)
    g.SetCoeffs(c)
    r = orc.CreateTriangularOrders(ref_opts(**opts), list(range(1, 33)), orders)
    r.SetCoeffs(c)
    x = np.ascontiguousarray(student_t_points(32, 150, nu=5, seed=9))
    lr, gr = r.KL(x)
    for lanes in (1, 2, 4, 8, 64):
        g.SetPointGroup(lanes)
        lf, gf = g.KLValueAndGrad(x)
        close(lf, lr, what=f"KL loss lanes={lanes}")
        close(gf, gr, what=f"KL grad lanes={lanes}")
    g.SetPointGroup(0)
    xl = torch.from_numpy(student_t_points(32, 20000, nu=5, seed=10)).cuda()
    lt, gt = g.KLValueAndGrad(xl)
    (_, _, lf_), trials = g.EvalSchedule()
    print(f"[point_group] tuned lanes fused={lf_} trials={trials[2]}", flush=True)
    assert lf_ in (1, 2, 4, 8, 64) and len(trials[2]) == 5 and all(t > 0 for t in trials[2])
    g.SetPointGroup(1)
    l1, g1 = g.KLValueAndGrad(xl)
    close(lt.cpu().numpy(), l1.cpu().numpy(), rtol=1e-12, what="tuned KL loss vs lane per point")
    close(gt.cpu().numpy(), g1.cpu().numpy(), rtol=1e-12, what="tuned KL grad vs lane per point")
