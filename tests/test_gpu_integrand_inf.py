"""The reference's infinite-integrand failure (MParT/MonotoneIntegrand.h:183-190):
g(df) = inf at a quadrature node makes the call fail with a runtime error.
The kernels record it in the map's status word; synchronous calls (numpy
inputs, host memory at the C ABI, the KL objective) raise, device calls leave
it to CheckIntegrand.  Each case is compared with the oracle, which raises
where the reference throws.

Map: one component in 1-D, order 2, Exp bijector, coefficient 1000 on the
normalized He_2, so df(y) ~ 893 y and exp(df) overflows for y > 0.795:
x = 1.0 meets an infinite g inside its quadrature, x = 0.5 does not."""
import ctypes

import numpy as np
import pytest

import mpart_amd as mt
from oracle import oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

COEFFS = np.array([0.0, 0.0, 1000.0])


def _pair(posFunc="Exp", contDeriv=True, coeffs=COEFFS):
    o = mt.MapOptions(posFuncType=mt.PosFuncTypes[posFunc], contDeriv=contDeriv)
    g = mt.CreateTotalOrderComponent(1, 2, o)
    g.SetCoeffs(coeffs)
    r = orc.CreateComponent(orc.options(posFuncType=posFunc, contDeriv=contDeriv), 1, 2)
    r.SetCoeffs(coeffs)
    return g, r


def _raises_like_oracle(fn_gpu, fn_ref):
    """Both raise the runtime error, or neither does (and then they agree)."""
    try:
        ref = fn_ref()
        ref_err = None
    except RuntimeError as e:
        ref, ref_err = None, e
    if ref_err is not None:
        with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
            fn_gpu()
        return None
    got = fn_gpu()
    return got, ref


def test_infinite_integrand_numpy_paths():
    print("[gpu-case] integrand inf: numpy paths", flush=True)
    g, r = _pair()
    ok = np.array([[0.5]])
    bad = np.array([[0.5, 1.0]])
    got, ref = _raises_like_oracle(lambda: g.Evaluate(ok), lambda: r.Evaluate(ok))
    assert abs(got[0, 0] - ref[0, 0]) <= 1e-10 * abs(ref[0, 0])
    g.CheckIntegrand()  # nothing recorded
    for name, fg, fr in [
        ("Evaluate", lambda: g.Evaluate(bad), lambda: r.Evaluate(bad)),
        ("CoeffGrad", lambda: g.CoeffGrad(bad, np.ones((1, 2))), lambda: r.CoeffGrad(bad, np.ones((1, 2)))),
        ("Gradient", lambda: g.Gradient(bad, np.ones((1, 2))), lambda: r.Gradient(bad, np.ones((1, 2)))),
    ]:
        print(f"[gpu-case] integrand inf: {name}", flush=True)
        assert _raises_like_oracle(fg, fr) is None, name
        g.CheckIntegrand()  # the failing call consumed the record
    # LogDeterminant evaluates g outside the integrand: inf, no error
    ld = g.LogDeterminant(np.array([[1.0]]))
    assert ld[0] == np.inf and r.LogDeterminant(np.array([[1.0]]))[0] == np.inf
    g.CheckIntegrand()
    # the KL objective on samples that reach the overflow
    obj = mt.CreateGaussianKLObjective(bad)
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        obj.TrainError(g)
    g.CheckIntegrand()
    # SoftPlus stays finite at 893
    gs, rs = _pair("SoftPlus")
    got, ref = _raises_like_oracle(lambda: gs.Evaluate(bad), lambda: rs.Evaluate(bad))
    assert np.all(np.abs(got - ref) <= 1e-10 * np.abs(ref))
    # discrete derivative: LogDeterminant integrates, so it fails
    gd, rd = _pair(contDeriv=False)
    assert _raises_like_oracle(lambda: gd.LogDeterminant(bad), lambda: rd.LogDeterminant(bad)) is None


def test_infinite_integrand_inverse_like_oracle():
    """Inverse's bracket search evaluates T where the integrand may overflow:
    the GPU raises exactly for the targets the oracle raises for.  Coefficient
    500: df ~ 446 y overflows for y > 1.59, T(1) ~ 1e191; the bracket [0, 1]
    holds the targets T(0.3) and 1e30, while 1e200 sends the search past 1.59."""
    print("[gpu-case] integrand inf: inverse", flush=True)
    g, r = _pair(coeffs=np.array([0.0, 0.0, 500.0]))
    t_small = r.Evaluate(np.array([[0.3]]))[0, 0]
    x1 = np.zeros((1, 1))  # a component's x1: its inputs, the last row the initial guess
    raised = []
    for tgt in [t_small, 1e30, 1e200]:
        res = _raises_like_oracle(lambda: g.Inverse(x1, np.array([[tgt]])),
                                  lambda: r.Inverse(x1, np.array([[tgt]])))
        raised.append(res is None)
        if res is not None:
            got, ref = res
            assert abs(got[0, 0] - ref[0, 0]) <= 1e-9, (tgt, got, ref)
        g.CheckIntegrand()
    assert raised == [False, False, True]


def test_infinite_integrand_device_calls_record():
    """Device calls do not fail by themselves; CheckIntegrand reports the record
    once and clears it.  The C ABI with host memory fails like the reference's
    host call."""
    print("[gpu-case] integrand inf: device record", flush=True)
    g, _ = _pair()
    xd = torch.tensor([[0.5, 1.0]], dtype=torch.float64, device="cuda")
    y = g.Evaluate(xd)  # returns
    assert torch.isinf(y[0, 1]) or torch.isnan(y[0, 1])
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        g.CheckIntegrand()
    g.CheckIntegrand()  # cleared
    # the fused KL pass records it too
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(3, dtype=torch.float64, device="cuda")
    g.KLValueAndGradImpl(xd, loss, grad)
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        g.CheckIntegrand()
    g.KLValueAndGradImpl(xd[:, :1], loss, grad)
    g.CheckIntegrand()
    # batched Evaluate: only the map whose points overflow has a record
    g2, _ = _pair()
    b = mt.MapBatch([g, g2])
    b.Evaluate([xd, xd[:, :1].contiguous()])
    with pytest.raises(RuntimeError, match="MonotoneIntegrand"):
        g.CheckIntegrand()
    g2.CheckIntegrand()
    # host memory straight through the C ABI
    xh = np.array([[0.5, 1.0]])
    yh = np.zeros((1, 2))
    v = lambda a: mt._CView(a.ctypes.data, a.shape[0], a.shape[1], a.strides[0] // 8, a.strides[1] // 8)
    rc = mt._lib.mpart_evaluate(g.handle, v(xh), v(yh), mt._stream())
    assert rc == -2 and "MonotoneIntegrand" in mt._lib.mpart_last_error().decode()
    rc = mt._lib.mpart_evaluate(g.handle, v(xh[:, :1]), v(yh[:, :1]), mt._stream())
    assert rc == 0
