"""TrainMap (src/TrainMap.cpp:168-216) through mpart_train_map: L-BFGS on the
host over fused KL passes on the device, a triangular map trained one
component at a time.

Checks: the returned error is the oracle's objective mean(0.5 |T|^2 - logdet)
at the trained coefficients (1e-10 relative); with tight tolerances the
gradient vanishes and the minimum agrees with torch.optim.LBFGS run through
the fused objective; the evaluation budget is respected; the reference's
loose default tolerances still decrease the objective.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream, banana_points
from mpart_amd.training import train_kl_lbfgs
from oracle import oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu-case] {name}", flush=True)


def _oracle_objective(dim, order, coeffs, x):
    r = orc.CreateTriangular(dim, dim, order, orc.options())
    r.SetCoeffs(coeffs)
    T = r.Evaluate(x)
    return float(np.mean(0.5 * np.sum(T * T, axis=0) - r.LogDeterminant(x)))


def _banana(n, dim, seed):
    return np.ascontiguousarray(banana_points(n, dim, seed=seed))


def test_train_triangular_tight_tolerances():
    _case("train_triangular")
    dim, order = 3, 2
    x = _banana(20000, dim, 3)
    xd = torch.from_numpy(x).cuda()
    m = mt.CreateTriangular(dim, dim, order, mt.MapOptions())
    obj = mt.CreateGaussianKLObjective(xd)
    e0 = obj.TrainError(m)
    opts = mt.TrainOptions(opt_ftol_rel=1e-14, opt_ftol_abs=1e-14, opt_xtol_rel=1e-12, opt_xtol_abs=1e-12,
                           opt_maxeval=2000)
    err = mt.TrainMap(m, obj, opts)
    ref = _oracle_objective(dim, order, m.CoeffMap(), x)
    assert abs(err - ref) <= 1e-10 * abs(ref), (err, ref)
    # TrainError adds the Gaussian normalization d log(2 pi) / 2
    assert abs(obj.TrainError(m) - (ref + 0.5 * dim * np.log(2 * np.pi))) <= 1e-10 * abs(ref) + 1e-12
    assert obj.TrainError(m) < e0
    _, g = m.KLValueAndGrad(xd)
    assert float(torch.linalg.norm(g)) < 1e-5
    # the same minimum through torch.optim.LBFGS on the fused objective
    m2 = mt.CreateTriangular(dim, dim, order, mt.MapOptions())
    _, hist = train_kl_lbfgs(m2, xd, steps=400)
    l2, _ = m2.KLValueAndGrad(xd)
    assert abs(float(l2) - err) <= 1e-7 * abs(err), (float(l2), err)
    print(f"[train] error {err:.12f} evaluations {m.last_train_evaluations}, torch L-BFGS {float(l2):.12f}", flush=True)


def test_train_component_and_budget():
    _case("train_component_budget")
    s = Stream(5)
    x = s.points(3, 5000)
    xd = torch.from_numpy(x).cuda()
    g = mt.CreateTotalOrderComponent(3, 3, mt.MapOptions())
    obj = mt.CreateGaussianKLObjective(xd, dim=1)  # the Gaussian's dimension: the map's output dimension
    e0 = obj.TrainError(g)
    err = mt.TrainMap(g, obj, mt.TrainOptions(opt_maxeval=6))
    assert g.last_train_evaluations <= 6
    assert obj.TrainError(g) < e0
    # default (reference) tolerances
    g2 = mt.CreateTotalOrderComponent(3, 3, mt.MapOptions())
    err2 = mt.TrainMap(g2, obj)
    assert err2 <= err + 1e-12 and g2.last_train_evaluations <= 1000
    r = orc.CreateComponent(orc.options(), 3, 3)
    r.SetCoeffs(g2.CoeffMap())
    T = r.Evaluate(x)[0]
    # a lone component is trained under the KL objective, whose value carries
    # the Gaussian normalization 0.5 d log(2 pi), d = 1 (src/TrainMap.cpp:184-199)
    ref = float(np.mean(0.5 * T * T - r.LogDeterminant(x))) + 0.5 * np.log(2.0 * np.pi)
    assert abs(err2 - ref) <= 1e-10 * abs(ref)
    assert abs(err2 - obj.TrainError(g2)) <= 1e-10 * abs(ref)


def test_train_rejects_bad_arguments():
    _case("train_errors")
    m = mt.CreateTriangular(2, 2, 2, mt.MapOptions())
    xd = torch.zeros((2, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(NotImplementedError):
        mt.TrainMap(m, mt.CreateGaussianKLObjective(xd), mt.TrainOptions(opt_alg="LN_COBYLA"))
    with pytest.raises(ValueError):
        mt.TrainMap(m, mt.CreateGaussianKLObjective(xd), mt.TrainOptions(opt_maxeval=0))
    with pytest.raises(ValueError):
        mt.TrainMap(m, mt.CreateGaussianKLObjective(torch.zeros((3, 10), dtype=torch.float64, device="cuda")))
