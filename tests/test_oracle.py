"""Pins the CPU oracle (oracle/mpart_oracle.cpp) before anything is compared with it.

Each check is a known answer from the reference's own test suite, a closed form,
or an independent computation (numpy's Hermite modules, Gauss-Legendre,
brute-force enumeration, finite differences) chosen so that a dropped term, a
wrong sign or index, or a transposed operand fails it.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as orc


# --------------------------------------------------------------------------
# Multi-index enumeration (src/MultiIndices/FixedMultiIndexSet.cpp:487-549)
@pytest.mark.parametrize("dim,order", [(1, 0), (1, 3), (2, 2), (2, 5), (3, 3), (6, 4), (8, 3)])
def test_total_order_matches_brute_force(dim, order):
    starts, dims, orders, maxdeg = orc.total_order(dim, order)
    # Brute force: lexicographic order, dim 0 outermost, last dim innermost.
    multis = [a for a in itertools.product(range(order + 1), repeat=dim) if sum(a) <= order]
    assert len(starts) - 1 == len(multis) == math.comb(dim + order, order)
    for t, a in enumerate(multis):
        nz = [(d, o) for d, o in enumerate(a) if o > 0]
        got = list(zip(dims[starts[t]:starts[t + 1]].tolist(), orders[starts[t]:starts[t + 1]].tolist()))
        assert got == nz, (t, a, got)
    # tests/MultiIndices/Test_MultiIndexSet.cpp:17-22: max degrees equal the order.
    assert maxdeg.tolist() == [order] * dim


def test_total_order_sizes_for_configs():
    # Term counts used by the configs: C3 components and C2.
    assert [len(orc.total_order(d, 3)[0]) - 1 for d in range(1, 9)] == [4, 10, 20, 35, 56, 84, 120, 165]
    s, d, o, _ = orc.total_order(6, 4)
    assert len(s) - 1 == 210 and len(d) == 504


# --------------------------------------------------------------------------
# 1-D bases
@pytest.mark.parametrize("x", [-2.3, -0.7, 0.0, 0.4, 1.9])
def test_prob_hermite_identities(x):
    # tests/Test_OrthogonalPolynomials.cpp:66-147 (unnormalized He_0..He_4).
    v, d, d2 = orc.basis_eval("ProbabilistHermite", x, 4, mode=2, normalize=False)
    he = [1.0, x, x * x - 1.0, x ** 3 - 3 * x, x ** 4 - 6 * x * x + 3]
    dhe = [0.0, 1.0, 2 * x, 3 * x * x - 3, 4 * x ** 3 - 12 * x]
    d2he = [0.0, 0.0, 2.0, 6 * x, 12 * x * x - 12]
    np.testing.assert_allclose(v, he, rtol=0, atol=1e-13)
    np.testing.assert_allclose(d, dhe, rtol=0, atol=1e-13)
    np.testing.assert_allclose(d2, d2he, rtol=0, atol=1e-13)
    # EvaluateAll and EvaluateDerivatives share the value recurrence.
    va = orc.basis_eval("ProbabilistHermite", x, 4, mode=0, normalize=False)[0]
    assert np.array_equal(va, v)


def test_prob_hermite_normalization_gauss_hermite():
    # basisNorm=true divides by sqrt(sqrt(2 pi) k!); the normalized family is
    # orthonormal under exp(-x^2/2) (textbook Gauss-Hermite check, cf.
    # tests/Test_OrthogonalPolynomials.cpp:11-63).
    xs, ws = np.polynomial.hermite_e.hermegauss(30)
    P = 8
    V = np.array([orc.basis_eval("ProbabilistHermite", x, P, mode=0, normalize=True)[0] for x in xs])
    G = (V * ws[:, None]).T @ V
    np.testing.assert_allclose(G, np.eye(P + 1), atol=1e-12)


@pytest.mark.parametrize("x", [-3.1, -1.0, 0.0, 0.3, 2.2])
def test_hermite_functions_closed_form(x):
    # phi_0 = 1, phi_1 = x, phi_{k+2} = psi_k with psi_k the orthonormal
    # physicist Hermite function (MParT/HermiteFunction.h:22-104,
    # tests/Test_HermiteFunctions.cpp:9-103).  Independent: numpy.polynomial.hermite.
    P = 7
    v, d, d2 = orc.basis_eval("HermiteFunctions", x, P, mode=2)
    va = orc.basis_eval("HermiteFunctions", x, P, mode=0)[0]
    psi, dpsi = [], []
    for k in range(P - 1):
        c = np.zeros(k + 1)
        c[k] = 1.0
        Hk = np.polynomial.hermite.hermval(x, c)
        dHk = np.polynomial.hermite.hermval(x, np.polynomial.hermite.hermder(c))
        s = (2.0 ** k * math.factorial(k) * math.sqrt(math.pi)) ** -0.5
        psi.append(s * Hk * math.exp(-0.5 * x * x))
        dpsi.append(s * (dHk - x * Hk) * math.exp(-0.5 * x * x))
    np.testing.assert_allclose(v, [1.0, x] + psi, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(va, [1.0, x] + psi, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(d, [0.0, 1.0] + dpsi, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(d2[2:], [-(2 * k + 1 - x * x) * psi[k] for k in range(P - 1)], rtol=1e-12, atol=1e-14)
    assert d2[0] == 0.0 and d2[1] == 0.0


# --------------------------------------------------------------------------
# Positive bijectors (tests/Test_PositiveBijectors.cpp:13-39)
@pytest.mark.parametrize("x", [-40.0, -3.0, -0.5, 0.0, 0.5, 3.0, 40.0])
def test_softplus_and_exp(x):
    g, dg, d2g = orc.posfunc_eval("SoftPlus", x)
    sig = 1.0 / (1.0 + math.exp(-x))
    # log(1 + exp(-|x|)) + max(x, 0): for x << 0 the sum 1 + exp(x) rounds to 1,
    # so the reference returns 0 there (absolute error below 1e-16).
    assert g == pytest.approx(math.log1p(math.exp(x)) if x < 30 else x, rel=1e-14, abs=1e-16)
    assert dg == pytest.approx(sig, rel=1e-14)
    assert d2g == pytest.approx(sig * (1 - sig), rel=1e-12, abs=1e-300)
    g, dg, d2g = orc.posfunc_eval("Exp", x)
    assert g == dg == d2g == math.exp(x)


# --------------------------------------------------------------------------
# Quadrature
def test_cc_rule_properties():
    # 3-point rule on [-1,1] is Simpson: weights 1/3, 4/3, 1/3.
    w, p = orc.cc_rule(3)
    np.testing.assert_allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15)
    np.testing.assert_array_equal(p, [-1.0, 0.0, 1.0])
    for L in range(1, 5):
        wc, pc = orc.cc_rule(2 ** L + 1)
        wf, pf = orc.cc_rule(2 ** (L + 1) + 1)
        # Nested points (tests/Test_Quadrature.cpp:24-42).
        assert np.array_equal(pc, pf[::2])
        # Exact for polynomials up to degree n-1.
        n = len(wf)
        for k in range(n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert np.dot(wf, pf ** k) == pytest.approx(exact, abs=1e-13)


@pytest.mark.parametrize("quad,level,maxSub,tol", [("AdaptiveSimpson", 0, 30, 1e-8),
                                                   ("AdaptiveClenshawCurtis", 2, 10, 1e-6),
                                                   ("AdaptiveClenshawCurtis", 1, 30, 1e-8)])
def test_adaptive_quadrature_known_answers(quad, level, maxSub, tol):
    # tests/Test_Quadrature.cpp:118-219 (ACC) and 223-318 (AS).
    val, _ = orc.quad_test(quad, level, maxSub, tol, tol, mode=0)
    assert val == pytest.approx(math.e - 1.0, rel=1e-4)
    assert abs(val - (math.e - 1.0)) < 50 * tol
    val, ne = orc.quad_test(quad, level, maxSub, tol, tol, mode=1)
    true = 0.5 + math.e - 1.0
    assert val == pytest.approx(true, rel=1e-4)
    assert ne < (150 if quad == "AdaptiveSimpson" else 300)


def test_adaptive_simpson_eval_count_formula():
    # Evaluations = 3 + 2 V (AS); with a smooth integrand the root interval is
    # accepted at the loosest tolerance: exactly 5 evaluations.
    _, ne = orc.quad_test("AdaptiveSimpson", 0, 30, 1.0, 1.0, mode=0)
    assert ne == 5
    _, ne = orc.quad_test("AdaptiveClenshawCurtis", 1, 30, 1.0, 1.0, mode=0)
    assert ne == 5  # 2 + 3 V with V = 1


# --------------------------------------------------------------------------
# Root finding (tests/Test_RootFinding.cpp:17-113)
def test_find_bracket_cases():
    sig = lambda v: 1.0 / (math.exp(-v) + 1)
    xlb, ylb, xub, yub, info = orc.find_bracket_test(0, -1.1, 0.0, 0.0, 2.0, 2.0)
    assert xlb < -1.1 < xub and ylb < -1.1 < yub and info == 0
    xlb, ylb, xub, yub, info = orc.find_bracket_test(1, sig(-0.5), 0.0, sig(0.0), 2.0, sig(2.0))
    assert xlb < -0.5 < xub and info == 0
    assert ylb == sig(xlb) and yub == sig(xub)
    *_, info = orc.find_bracket_test(3, 0.0, 0.0, -1.0, 2.0, -1.0)
    assert info == -1


@pytest.mark.parametrize("fn,x0", [(0, 0.0), (0, 1.0), (1, 0.0), (1, 1.0), (2, -5.0), (2, 5.0)])
def test_itp_inverse_cases(fn, x0):
    f = [lambda v: v, lambda v: 1 / (math.exp(-v) + 1),
         lambda v: sum(a / (math.exp(-b * (v - c)) + 1) for a, b, c in [(1, 2, 1), (3, 0.5, 0), (0.5, 1.5, -1)])][fn]
    x, info = orc.itp_test(fn, f(0.5), x0, 1e-5, 1e-5)
    assert x == pytest.approx(0.5, rel=2e-5) and info == 0


def test_itp_flat_returns_nan():
    x, info = orc.itp_test(3, 0.0, 0.0, 1e-5, 1e-5)
    assert math.isnan(x) and info == -2


# --------------------------------------------------------------------------
# Monotone components: closed forms (tests/Test_MonotoneComponent.cpp:360-557)
def _unnorm_exp_as(maxSub, tol):
    return orc.options(posFuncType="Exp", quadType="AdaptiveSimpson", basisNorm=False, quadMaxSub=maxSub,
                       quadAbsTol=tol, quadRelTol=tol)


def test_component_affine_closed_form():
    comp = orc.CreateComponent(_unnorm_exp_as(3, 1e-7), 1, 1)
    comp.SetCoeffs(np.array([1.0, 1.0]))
    x = np.array([[(i / (20 - 0.5)) * 4.0 + 2.0 for i in range(20)]])
    np.testing.assert_allclose(comp.Evaluate(x)[0], 1 + math.e * x[0], rtol=1e-7)


def test_component_quadratic_closed_form_and_logdet():
    comp = orc.CreateComponent(_unnorm_exp_as(30, 1e-7), 1, 2)
    comp.SetCoeffs(np.array([1.5, 1.0, 0.5]))  # f = 1 + x + 0.5 x^2
    x = np.linspace(-2.0, 2.0, 20)[None, :]
    np.testing.assert_allclose(comp.Evaluate(x)[0], 1 + math.e * (np.exp(x[0]) - 1), rtol=1e-7)
    # d/dx T = exp(1 + x) exactly for the continuous derivative.
    np.testing.assert_allclose(comp.LogDeterminant(x), 1 + x[0], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("coeffs,deg", [([1.0, 1.0], 1), ([1.5, 1.0, 0.5], 2)])
def test_component_inverse_round_trip(coeffs, deg):
    comp = orc.CreateComponent(_unnorm_exp_as(30, 1e-7), 1, deg)
    comp.SetCoeffs(np.array(coeffs))
    x = np.array([[(i / 19.0) * 4.0 - 2.0 for i in range(20)]])
    y = comp.Evaluate(x)
    np.testing.assert_allclose(comp.Inverse(x, y)[0], x[0], rtol=1e-6, atol=1e-6)
    # Same x, multiple ys (Test_MonotoneComponent.cpp:527-557).
    np.testing.assert_allclose(comp.Inverse(np.array([[0.0]]), y)[0], x[0], rtol=1e-6, atol=1e-6)


def test_inverse_nan_input_gives_nan():
    comp = orc.CreateComponent(orc.options(), 2, 2)
    comp.SetCoeffs(np.linspace(-0.2, 0.3, comp.numCoeffs))
    xs = np.array([[np.nan, 0.1], [0.0, 0.0]])
    out = comp.Inverse(xs, np.array([0.5, 0.5]))
    assert math.isnan(out[0, 0]) and math.isfinite(out[0, 1])


# --------------------------------------------------------------------------
# Derivatives against central finite differences (tests/Test_MonotoneComponent.cpp:561-869)
def _rand_component(opts, dim, order, seed):
    rng = np.random.default_rng(seed)
    comp = orc.CreateComponent(opts, dim, order)
    c = 0.2 * rng.standard_normal(comp.numCoeffs)
    comp.SetCoeffs(c)
    x = rng.standard_normal((dim, 25))
    return comp, c, x


OPTS = [orc.options(),
        orc.options(basisType="HermiteFunctions", posFuncType="Exp"),
        orc.options(quadType="AdaptiveClenshawCurtis", quadAbsTol=1e-8, quadRelTol=1e-8),
        orc.options(posFuncType="Exp", quadType="AdaptiveClenshawCurtis", basisNorm=False, nugget=0.1)]


@pytest.mark.parametrize("oi", range(len(OPTS)))
def test_coeff_grad_and_logdet_coeff_grad_fd(oi):
    comp, c, x = _rand_component(OPTS[oi], 3, 3, 10 + oi)
    sens = np.ones((1, x.shape[1]))
    G = comp.CoeffGrad(x, sens)
    LG = comp.LogDeterminantCoeffGrad(x)
    h = 1e-6
    for k in range(comp.numCoeffs):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        comp.SetCoeffs(cp)
        ep, lp = comp.Evaluate(x)[0], comp.LogDeterminant(x)
        comp.SetCoeffs(cm)
        em, lm = comp.Evaluate(x)[0], comp.LogDeterminant(x)
        np.testing.assert_allclose(G[k], (ep - em) / (2 * h), rtol=1e-4, atol=1e-6)
        np.testing.assert_allclose(LG[k], (lp - lm) / (2 * h), rtol=1e-4, atol=1e-6)
    comp.SetCoeffs(c)
    # Scaling by the sensitivity is exact (Test_MonotoneComponent.cpp:825-840).
    s2 = np.linspace(-1, 2, x.shape[1])[None, :]
    assert np.array_equal(comp.CoeffGrad(x, s2), G * s2)


@pytest.mark.parametrize("oi", range(len(OPTS)))
def test_input_gradients_fd(oi):
    comp, c, x = _rand_component(OPTS[oi], 3, 3, 20 + oi)
    sens = np.ones((1, x.shape[1]))
    G = comp.Gradient(x, sens)
    LG = comp.LogDeterminantInputGrad(x)
    h = 1e-6
    # Reference quirk: Evaluate drops the nugget (MonotoneComponent.h:287) while
    # InputJacobian integrates x_d (g + nugget) (MonotoneIntegrand.h:181,201),
    # so d/dx_d of Gradient exceeds the finite difference of Evaluate by the nugget.
    nug = OPTS[oi].nugget
    for d in range(3):
        xp, xm = x.copy(), x.copy()
        xp[d] += h
        xm[d] -= h
        fd = (comp.Evaluate(xp)[0] - comp.Evaluate(xm)[0]) / (2 * h) + (nug if d == 2 else 0.0)
        np.testing.assert_allclose(G[d], fd, rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(LG[d], (comp.LogDeterminant(xp) - comp.LogDeterminant(xm)) / (2 * h),
                                   rtol=1e-4, atol=1e-6)


def test_discrete_derivative_mode_consistent():
    # contDeriv=false: log det of the quadrature approximation; close to the
    # continuous one and its coefficient gradient matches finite differences.
    oc = orc.options()
    od = orc.options(contDeriv=False)
    cc, c, x = _rand_component(oc, 2, 3, 5)
    cd = orc.CreateComponent(od, 2, 3)
    cd.SetCoeffs(c)
    np.testing.assert_allclose(cd.LogDeterminant(x), cc.LogDeterminant(x), atol=1e-4)
    LG = cd.LogDeterminantCoeffGrad(x)
    h = 1e-6
    for k in range(cd.numCoeffs):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        cd.SetCoeffs(cp)
        lp = cd.LogDeterminant(x)
        cd.SetCoeffs(cm)
        lm = cd.LogDeterminant(x)
        np.testing.assert_allclose(LG[k], (lp - lm) / (2 * h), rtol=1e-3, atol=1e-5)
    cd.SetCoeffs(c)


# --------------------------------------------------------------------------
# Triangular maps (tests/Test_TriangularMap.cpp:10-228)
def test_triangular_block_consistency():
    opts = orc.options()
    tri = orc.CreateTriangular(3, 3, 2, opts)
    assert tri.numCoeffs == sum(math.comb(d + 2, 2) for d in (1, 2, 3))
    rng = np.random.default_rng(3)
    c = 0.1 * rng.standard_normal(tri.numCoeffs)
    tri.SetCoeffs(c)
    x = rng.standard_normal((3, 40))
    Y = tri.Evaluate(x)
    LD = tri.LogDeterminant(x)
    off = 0
    ld = np.zeros(x.shape[1])
    for d in (1, 2, 3):
        comp = orc.CreateComponent(opts, d, 2)
        comp.SetCoeffs(c[off:off + comp.numCoeffs])
        off += comp.numCoeffs
        np.testing.assert_array_equal(Y[d - 1], comp.Evaluate(x[:d])[0])
        ld += comp.LogDeterminant(x[:d])
    np.testing.assert_array_equal(LD, ld)
    # Inverse round trip.
    xi = tri.Inverse(np.zeros((0, x.shape[1])), Y)
    np.testing.assert_allclose(xi, x, atol=1e-5)


def test_coeff_count_rectangular():
    # bindings/python/tests/test_MapFactory.py:69-71: CreateTriangular(2,2,2) has 2+7? -> 3 + 6 coeffs.
    tri = orc.CreateTriangular(2, 2, 2, orc.options())
    assert tri.numCoeffs == 3 + 6
    tri = orc.CreateTriangular(3, 2, 2, orc.options())
    assert tri.numCoeffs == 6 + 10


def test_kl_gradient_fd_and_constant():
    opts = orc.options()
    tri = orc.CreateTriangular(3, 3, 2, opts)
    rng = np.random.default_rng(7)
    c = 0.1 * rng.standard_normal(tri.numCoeffs)
    tri.SetCoeffs(c)
    x = rng.standard_normal((3, 60))
    loss, grad = tri.KL(x)
    loss_c, grad_c = tri.KL(x, with_const=True)
    # KLObjective adds d/2 log(2 pi); the gradients are identical
    # (tests/Test_MapObjective.cpp:119-141).
    assert loss_c - loss == pytest.approx(1.5 * math.log(2 * math.pi), rel=1e-12)
    assert np.array_equal(grad, grad_c)
    h = 1e-6
    for k in range(tri.numCoeffs):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        tri.SetCoeffs(cp)
        lp, _ = tri.KL(x)
        tri.SetCoeffs(cm)
        lm, _ = tri.KL(x)
        assert grad[k] == pytest.approx((lp - lm) / (2 * h), rel=1e-4, abs=1e-6)
    tri.SetCoeffs(c)
    # Loss equals mean(0.5 |T|^2 - logdet) computed from Evaluate/LogDeterminant.
    T = tri.Evaluate(x)
    assert loss == pytest.approx(np.mean(0.5 * (T ** 2).sum(0) - tri.LogDeterminant(x)), rel=1e-13)


# --------------------------------------------------------------------------
# Infinite integrand (MParT/MonotoneIntegrand.h:183-190): g(df) = inf at a
# quadrature node raises; LogDeterminant, which evaluates g outside the
# integrand (ContinuousDerivative), does not.  One component in 1-D, order 2,
# coefficient 1000 on the normalized He_2: df(y) = 2000 y / sqrt(2 sqrt(2 pi))
# ~ 893 y, so exp(df) overflows for y > 0.795 (x = 1.0 reaches it at t > 0.795;
# x = 0.5 never does).
def _inf_component(posFunc="Exp", **kw):
    c = orc.CreateComponent(orc.options(posFuncType=posFunc, **kw), 1, 2)
    c.SetCoeffs(np.array([0.0, 0.0, 1000.0]))
    return c


def test_infinite_integrand_raises():
    c = _inf_component()
    y = c.Evaluate(np.array([[0.5]]))
    assert np.isfinite(y).all() and y[0, 0] > 1e150  # ~ e^446.5 / 1786
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        c.Evaluate(np.array([[0.5, 1.0]]))
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        c.CoeffGrad(np.array([[1.0]]), np.ones((1, 1)))
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        c.Gradient(np.array([[1.0]]), np.ones((1, 1)))
    # LogDeterminant does not integrate: g(df(1)) = inf gives log = inf, no error
    assert c.LogDeterminant(np.array([[1.0]]))[0] == np.inf
    # SoftPlus(893) is finite: no error
    assert np.isfinite(_inf_component("SoftPlus").Evaluate(np.array([[1.0]]))).all()
    # the discrete derivative integrates: its LogDeterminant raises
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        _inf_component(contDeriv=False).LogDeterminant(np.array([[1.0]]))
