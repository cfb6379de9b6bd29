"""Pins the oracle's remaining factory keys before the GPU is compared with it:
PhysicistHermite (MParT/OrthogonalPolynomial.h:313-326), LinearizedBasis
(MParT/LinearizedBasis.h) and the fixed Clenshaw-Curtis rule
(MParT/Quadrature.h:212-307).

Checks are closed forms (textbook Hermite polynomials, exact polynomial
integrals), an independent orthonormality computation (numpy's Gauss-Hermite
rule for the weight exp(-x^2)) and properties that determine the
linearization uniquely (continuity of value and slope at the bounds,
vanishing second differences outside, identity inside).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc


@pytest.mark.parametrize("x", [-2.3, -0.7, 0.0, 0.4, 1.9])
def test_physicist_hermite_closed_forms(x):
    v, d, d2 = orc.basis_eval("PhysicistHermite", x, 4, mode=2, normalize=False)
    H = [1.0, 2 * x, 4 * x * x - 2, 8 * x ** 3 - 12 * x, 16 * x ** 4 - 48 * x * x + 12]
    np.testing.assert_allclose(v, H, rtol=0, atol=1e-12)
    # H_k' = 2k H_{k-1}, H_k'' = 4k(k-1) H_{k-2}
    np.testing.assert_allclose(d, [0.0] + [2 * k * H[k - 1] for k in range(1, 5)], rtol=0, atol=1e-12)
    np.testing.assert_allclose(d2, [0.0, 0.0] + [4 * k * (k - 1) * H[k - 2] for k in range(2, 5)], rtol=0,
                               atol=1e-12)


def test_physicist_hermite_normalization():
    # basisNorm divides by sqrt(sqrt(pi) 2^k k!): orthonormal under exp(-x^2).
    xs, ws = np.polynomial.hermite.hermgauss(30)
    P = 8
    V = np.array([orc.basis_eval("PhysicistHermite", x, P, mode=0, normalize=True)[0] for x in xs])
    np.testing.assert_allclose((V * ws[:, None]).T @ V, np.eye(P + 1), atol=1e-12)


@pytest.mark.parametrize("basis", ["ProbabilistHermite", "PhysicistHermite", "HermiteFunctions"])
def test_linearized_basis_properties(basis):
    lb, ub, P = -1.2, 0.9, 5
    raw = lambda x, m: orc.basis_eval(basis, x, P, mode=m)
    lin = lambda x, m: orc.basis_eval_lin(basis, x, P, lb, ub, mode=m)
    # inside: the basis itself, bit for bit
    for x in (-1.2, -0.3, 0.0, 0.9):
        for a, b in zip(lin(x, 2), raw(x, 2)):
            assert np.array_equal(a, b)
    for bound, side in ((lb, -1.0), (ub, 1.0)):
        xs = bound + side * np.array([0.5, 1.0, 1.5, 4.0])
        V = np.array([lin(x, 2)[0] for x in xs])
        D = np.array([lin(x, 2)[1] for x in xs])
        D2 = np.array([lin(x, 2)[2] for x in xs])
        # linear outside: equal second differences vanish, slope constant, no curvature
        np.testing.assert_allclose(V[0] - 2 * V[1] + V[2], 0.0, atol=1e-12)
        assert np.all(D2 == 0.0)
        for k in range(1, 4):
            np.testing.assert_allclose(D[k], D[0], rtol=1e-14, atol=1e-14)
        # value and slope continuous at the bound (limit of the outside line)
        vb, db, _ = raw(bound, 2)
        np.testing.assert_allclose(D[0], db, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(V[0] - D[0] * (xs[0] - bound), vb, rtol=1e-12, atol=1e-13)
        # EvaluateAll agrees with EvaluateDerivatives' values
        np.testing.assert_allclose(lin(xs[1], 0)[0], V[1], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9, 17])
def test_fixed_clenshaw_curtis_exact_on_polynomials(n):
    # ClenshawCurtisQuadrature(n) on [0, 1] integrates t^k exactly for k < n
    # (k <= n for odd n, by symmetry).
    pts, wts = np.zeros(n), np.zeros(n)
    orc.lib().orc_cc_rule(n, orc._dp(wts), orc._dp(pts))
    t = 0.5 * (1.0 + pts)
    for k in range(n + (n % 2)):
        np.testing.assert_allclose(np.sum(0.5 * wts * t ** k), 1.0 / (k + 1), rtol=1e-13)
    r, ne = orc.quad_test("ClenshawCurtis", n, 0, 0.0, 0.0, 0)
    assert ne == n
    assert abs(r - (math.e - 1.0)) < (1e-12 if n >= 9 else 0.15)  # n = 2 is the trapezoid rule


def test_component_with_fixed_cc_and_linearized_basis():
    # T(x) = f(0) + x exp(c1) for f = c0 + c1 x (un-normalized, Exp): the
    # integrand is constant, so every rule is exact; a linearized basis with
    # bounds containing every evaluation point changes nothing.
    for quad in ("ClenshawCurtis", "AdaptiveSimpson"):
        for lbub in ((-math.inf, math.inf), (-10.0, 10.0)):
            o = orc.options(posFuncType="Exp", quadType=quad, basisNorm=False, quadPts=7,
                            basisLB=lbub[0], basisUB=lbub[1])
            comp = orc.CreateComponent(o, 1, 1)
            comp.SetCoeffs(np.array([1.0, 1.0]))
            x = np.linspace(-2.0, 2.0, 9)[None, :]
            np.testing.assert_allclose(comp.Evaluate(x)[0], 1 + math.e * x[0], rtol=1e-13)
    # linearized with x beyond ub: f = c0 + c1 He_1 + c2 He_2, He_2 linear past ub
    o = orc.options(posFuncType="Exp", quadType="ClenshawCurtis", basisNorm=False, quadPts=9,
                    basisLB=-1.0, basisUB=1.0)
    comp = orc.CreateComponent(o, 1, 2)
    comp.SetCoeffs(np.array([0.0, 0.0, 0.5]))  # f = 0.5 He_2 = 0.5 (x^2 - 1) inside
    # outside [-1, 1]: He_2 = 0 + 2 (x - 1) for x > 1, so d f/dx = 1 and T(x) = f(0) + int_0^1 x e^{d_x f(tx)} dt
    x = np.array([[3.0]])
    # d_x f(y) = y for |y| <= 1, 1 for y > 1: T(3) = -0.5 + int_0^{3} e^{min(y,1)} dy
    exact = -0.5 + (math.e - 1.0) + 2.0 * math.e
    assert abs(comp.Evaluate(x)[0, 0] - exact) < 0.02  # a 9-point rule on a kinked integrand
