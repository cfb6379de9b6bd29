"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the
same seeded inputs.

Tolerances (FP64): Evaluate / LogDeterminant / gradients agree to 1e-10
relative -- per element |gpu - ref| <= 1e-10 |ref| + 16 eps max|ref|_row, a
relative bound with a floor of a few ulps of the row's scale for entries that
cancel to ~0; Inverse
agrees to 1e-9 absolute.  The adaptive trees are compared through the
per-point integrand-evaluation counts.
"""
import math

import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream, banana_points
from oracle import oracle as orc
from parity import check

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _print_case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def close(gpu, ref, rtol=RTOL, what="", ulps=16.0):
    """Elementwise bound |gpu - ref| <= rtol |ref| + ulps eps max|ref_row|
    (parity.check: the floor is a few ulps of the row's scale and matters only
    where |ref| cancels below 1e-5 of the row maximum).  Returns worst/bound."""
    return check(gpu, ref, rtol=rtol, ulps=ulps, what=what)["worst"]


def ref_opts(**kw):
    return orc.options(**kw)


def gpu_opts(**kw):
    o = mt.MapOptions()
    for k, v in kw.items():
        if k == "basisType":
            v = mt.BasisTypes[v]
        elif k == "posFuncType":
            v = mt.PosFuncTypes[v]
        elif k == "quadType":
            v = mt.QuadTypes[v]
        setattr(o, k, v)
    return o


CASES = [
    dict(),
    dict(posFuncType="Exp"),
    dict(basisType="HermiteFunctions", posFuncType="Exp"),
    dict(basisType="HermiteFunctions"),
    dict(quadType="AdaptiveClenshawCurtis", quadAbsTol=1e-8, quadRelTol=1e-8),
    dict(quadType="AdaptiveClenshawCurtis", quadPts=9, posFuncType="Exp"),
    dict(basisNorm=False, nugget=0.05),
    dict(quadMaxSub=6, quadAbsTol=1e-10, quadRelTol=1e-10),
    # the remaining factory keys (MParT/MapFactory.h:218-245, src/MapFactoryImpl1-18.cpp)
    dict(basisType="PhysicistHermite"),
    dict(basisType="PhysicistHermite", posFuncType="Exp", quadType="AdaptiveClenshawCurtis", quadAbsTol=1e-8,
         quadRelTol=1e-8),
    dict(quadType="ClenshawCurtis", quadPts=9),
    dict(basisType="HermiteFunctions", posFuncType="Exp", quadType="ClenshawCurtis", quadPts=6),
    dict(basisLB=-1.5, basisUB=1.5),
    dict(basisType="HermiteFunctions", basisLB=-1.0, basisUB=2.0, quadType="AdaptiveClenshawCurtis",
         quadAbsTol=1e-8, quadRelTol=1e-8),
    dict(basisType="PhysicistHermite", basisLB=0.5, basisUB=3.0, quadType="ClenshawCurtis", quadPts=7),
]


def _pair(case, dim, order, seed, n=800, triangular=False, coeff_scale=0.2):
    s = Stream(seed)
    if triangular:
        g = mt.CreateTriangular(dim, dim, order, gpu_opts(**case))
        r = orc.CreateTriangular(dim, dim, order, ref_opts(**case))
    else:
        g = mt.CreateTotalOrderComponent(dim, order, gpu_opts(**case))
        r = orc.CreateComponent(ref_opts(**case), dim, order)
    x = s.points(dim, n)
    c = coeff_scale * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    return g, r, x, c


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_component_forward_and_trees(ci):
    _print_case("component_forward", case=CASES[ci])
    g, r, x, _ = _pair(CASES[ci], 3, 3, 100 + ci)
    yg, neg = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    assert np.array_equal(neg, ner), f"tree mismatch at {np.count_nonzero(neg != ner)} points"
    close(yg, yr, what="Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_component_gradients(ci):
    _print_case("component_gradients", case=CASES[ci])
    g, r, x, _ = _pair(CASES[ci], 3, 3, 200 + ci, n=400)
    sens = Stream(7).normal(x.shape[1])[None, :]
    close(g.CoeffGrad(x, sens), r.CoeffGrad(x, sens), what="CoeffGrad")
    close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad")
    close(g.Gradient(x, sens), r.Gradient(x, sens), what="Gradient")
    close(g.LogDeterminantInputGrad(x), r.LogDeterminantInputGrad(x), what="LogDeterminantInputGrad")


@pytest.mark.parametrize("ci", [0, 2, 4])
def test_triangular_all_methods(ci):
    _print_case("triangular", case=CASES[ci])
    g, r, x, _ = _pair(CASES[ci], 4, 3, 300 + ci, n=500, triangular=True, coeff_scale=0.1)
    yg, neg = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    assert np.array_equal(neg, ner)
    close(yg, yr, what="Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")
    sens = Stream(8).points(4, x.shape[1])
    close(g.CoeffGrad(x, sens), r.CoeffGrad(x, sens), what="CoeffGrad")
    close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad")
    close(g.Gradient(x, sens), r.Gradient(x, sens), what="Gradient")
    close(g.LogDeterminantInputGrad(x), r.LogDeterminantInputGrad(x), what="LogDeterminantInputGrad")
    # Per-component launches give the same bits as the batched launch.
    g.SetPlainLaunches(True)
    assert np.array_equal(g.Evaluate(x), yg)
    g.SetPlainLaunches(False)


@pytest.mark.parametrize("ci", [0, 1, 2, 4])
def test_kl_fused_and_plain(ci):
    _print_case("kl", case=CASES[ci])
    g, r, x, _ = _pair(CASES[ci], 4, 3, 400 + ci, n=700, triangular=True, coeff_scale=0.1)
    lr, gr = r.KL(x)
    lf, gf = g.KLValueAndGrad(x)
    lp, gp = g.KLValueAndGrad(x, plain=True)
    close(lf, lr, what="KL loss fused")
    close(gf, gr, what="KL grad fused")
    close(lp, lr, what="KL loss plain")
    close(gp, gr, what="KL grad plain")


def test_vjp_matches_materialized():
    torch = pytest.importorskip("torch")
    _print_case("vjp")
    g, r, x, _ = _pair({}, 3, 3, 500, n=600, triangular=True, coeff_scale=0.1)
    s = Stream(9)
    sens = s.points(3, x.shape[1])
    ls = s.normal(x.shape[1])
    ref = r.CoeffGrad(x, sens).sum(1) + (r.LogDeterminantCoeffGrad(x) * ls[None, :]).sum(1)
    xt = torch.from_numpy(x).cuda()
    out = g.CoeffVJP(xt, torch.from_numpy(sens).cuda(), torch.from_numpy(ls).cuda())
    close(out.cpu().numpy(), ref, what="CoeffVJP")


@pytest.mark.parametrize("ci", [0, 2, 4])
def test_component_inverse(ci):
    _print_case("inverse", case=CASES[ci])
    g, r, x, _ = _pair(CASES[ci], 3, 3, 600 + ci, n=300)
    y = r.Evaluate(x)
    # r ~ N(0,1) targets (config C2's Inverse): 1e-9 absolute.
    tgt = Stream(61).normal(x.shape[1])[None, :]
    xi_g = g.Inverse(x, tgt)
    xi_r = r.Inverse(x, tgt)
    fin = np.isfinite(xi_r)
    assert np.array_equal(fin, np.isfinite(xi_g))
    assert np.max(np.abs(xi_g[fin] - xi_r[fin])) <= 1e-9
    # Targets equal to T at the initial guess: the reference's FindBracket picks
    # its search direction from f(x0) > yd (RootFinding.h:35), a last-ulp
    # comparison here, so implementations agree only to the xtol scale
    # (the oracle itself moves by up to 4e-6 under a 1e-15 coefficient change).
    xs_g = g.Inverse(x, y)[0]
    xs_r = r.Inverse(x, y)[0]
    assert np.max(np.abs(xs_g - xs_r)) <= 1e-5
    np.testing.assert_allclose(xs_g, x[-1], atol=1e-5)


def test_triangular_inverse_round_trip():
    _print_case("triangular_inverse")
    g, r, x, _ = _pair({}, 3, 2, 700, n=300, triangular=True, coeff_scale=0.1)
    y = r.Evaluate(x)
    # Initial guesses are 0 here, so the self-target ambiguity above does not arise.
    xg = g.Inverse(np.zeros((0, x.shape[1])), y)
    xr = r.Inverse(np.zeros((0, x.shape[1])), y)
    assert np.max(np.abs(xg - xr)) <= 1e-9
    np.testing.assert_allclose(xg, x, atol=1e-5)


def test_inverse_nan_and_shapes():
    _print_case("inverse_nan")
    g, r, x, _ = _pair({}, 2, 2, 800, n=4)
    x = x.copy()
    x[0, 1] = np.nan
    out = g.Inverse(x, np.zeros(4))
    assert math.isnan(out[0, 1]) and np.isfinite(out[0, [0, 2, 3]]).all()
    with pytest.raises(ValueError):
        g.Inverse(x[:, :3], np.zeros(4))
    with pytest.raises(ValueError):
        g.Inverse(x, np.zeros(4), xtol=0.0, ytol=0.0)


def test_closed_form_components_on_gpu():
    # tests/Test_MonotoneComponent.cpp:360-440 (un-normalized ProbHermite, Exp).
    _print_case("closed_forms")
    o = gpu_opts(posFuncType="Exp", basisNorm=False, quadMaxSub=30, quadAbsTol=1e-7, quadRelTol=1e-7)
    c = mt.CreateTotalOrderComponent(1, 2, o)
    c.SetCoeffs(np.array([1.5, 1.0, 0.5]))
    x = np.linspace(-2, 2, 20)[None, :]
    np.testing.assert_allclose(c.Evaluate(x)[0], 1 + math.e * (np.exp(x[0]) - 1), rtol=1e-7)
    np.testing.assert_allclose(c.LogDeterminant(x), 1 + x[0], rtol=1e-12, atol=1e-14)


def test_strided_and_empty_inputs():
    torch = pytest.importorskip("torch")
    _print_case("strided")
    g, r, x, _ = _pair({}, 3, 2, 900, n=257, triangular=True, coeff_scale=0.1)
    ref = g.Evaluate(x)
    # point-major storage (x.T contiguous) viewed as D x N with rowStride 1
    xt = torch.from_numpy(np.ascontiguousarray(x.T)).cuda().T
    out = torch.zeros((3, x.shape[1]), dtype=torch.float64, device="cuda")
    g.EvaluateImpl(xt, out)
    assert np.array_equal(out.cpu().numpy(), ref)
    empty = torch.zeros((3, 0), dtype=torch.float64, device="cuda")
    out0 = torch.zeros((3, 0), dtype=torch.float64, device="cuda")
    g.EvaluateImpl(empty, out0)
    with pytest.raises(ValueError):
        g.Evaluate(x[:2])


def test_c1_full_size():
    # Config 1: 2-D triangular, order 3, ProbHermite/SoftPlus/AS, N = 10,000,
    # x ~ N(0, I_2) and coefficients ~ N(0, 0.1^2) from one seed-0 stream.
    _print_case("c1")
    s = Stream(0)
    x = s.points(2, 10000)
    g = mt.CreateTriangular(2, 2, 3, mt.MapOptions())
    r = orc.CreateTriangular(2, 2, 3, orc.options())
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    yg, neg = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    print(f"[c1] tree mismatches {np.count_nonzero(neg != ner)} / {neg.size}", flush=True)
    assert np.count_nonzero(neg != ner) == 0
    close(yg, yr, what="C1 Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="C1 LogDeterminant")


@pytest.mark.parametrize("variant", ["banana_zero_coeffs", "normal_data"])
def test_c3_sample_kl(variant):
    # Config 3 map on a sample: the literal banana inputs at the factory's zero
    # coefficients, and N(0, I_8) inputs with 0.1 N(0,1) coefficients (the
    # banana recursion explodes, so nonzero coefficients drive SoftPlus to
    # underflow and the log-determinant to -inf, as in the reference).
    _print_case("c3_sample", variant=variant)
    if variant == "banana_zero_coeffs":
        x = banana_points(3000)
        c = np.zeros(494)
    else:
        s = Stream(3)
        x = s.points(8, 3000)
        c = 0.1 * s.normal(494)
    g = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    r = orc.CreateTriangular(8, 8, 3, orc.options())
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    lr, gr = r.KL(x)
    lf, gf = g.KLValueAndGrad(x)
    lp, gp = g.KLValueAndGrad(x, plain=True)
    close(lf, lr, what="C3 loss")
    close(gf, gr, what="C3 grad")
    close(lp, lr, what="C3 loss plain")
    close(gp, gr, what="C3 grad plain")


DISCRETE_CASES = [
    dict(contDeriv=False),
    dict(contDeriv=False, posFuncType="Exp", basisType="HermiteFunctions"),
    dict(contDeriv=False, basisNorm=False, nugget=0.05),
    dict(contDeriv=False, quadType="AdaptiveClenshawCurtis", quadAbsTol=1e-8, quadRelTol=1e-8),
]


@pytest.mark.parametrize("ci", range(len(DISCRETE_CASES)))
def test_discrete_derivative_mode(ci):
    # contDeriv = false: LogDeterminant from DiscreteDerivative and its
    # coefficient gradient from DiscreteMixedJacobian (MonotoneComponent.h:586-638, 928-980).
    case = DISCRETE_CASES[ci]
    _print_case("discrete", case=case)
    g, r, x, _ = _pair(case, 3, 3, 950 + ci, n=600, triangular=True, coeff_scale=0.1)
    close(g.Evaluate(x), r.Evaluate(x), what="Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant (discrete)")
    close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad (discrete)")
    g.SetPlainLaunches(True)
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant (discrete, per component)")
    g.SetPlainLaunches(False)
    # the KL objective in one fused launch (the generic kernel with the
    # discrete derivative), with each schedule it can take, and the plain path
    lr, gr = r.KL(x)
    for lanes in (1, 2):
        g.SetPointGroup(lanes)
        lf, gf = g.KLValueAndGrad(x)
        close(lf, lr, what=f"KL loss (discrete, fused, schedule {lanes})")
        close(gf, gr, what=f"KL grad (discrete, fused, schedule {lanes})")
    g.SetPointGroup(0)
    lp, gp = g.KLValueAndGrad(x, plain=True)
    close(lp, lr, what="KL loss (discrete, plain)")
    close(gp, gr, what="KL grad (discrete, plain)")
    # the coefficient VJP of the torch backward: sens . dT/dw + lsens . dlogdet/dw
    torch = pytest.importorskip("torch")
    s = Stream(990 + ci)
    sens = s.normal(3 * x.shape[1]).reshape(3, x.shape[1])
    ls = s.normal(x.shape[1])
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    v = g.CoeffVJP(xd, torch.from_numpy(sens).cuda(), torch.from_numpy(ls).cuda()).cpu().numpy()
    ref = r.CoeffGrad(np.ascontiguousarray(x), sens).sum(axis=1) + r.LogDeterminantCoeffGrad(np.ascontiguousarray(x)) @ ls
    close(v, ref, what="coefficient VJP (discrete)")
    with pytest.raises(ValueError):
        g.LogDeterminantInputGrad(x)
