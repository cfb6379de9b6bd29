"""Fused input gradient of the KL objective against the oracle and torch.

The fused launch forms, per point, grad_x [0.5 |T(x)|^2 - log det grad T(x)]
= Gradient(x, sens = T(x)) - LogDeterminantInputGrad(x), the combination the
reference's torch wrapper builds from two separate calls
(bindings/python/package/torch_helpers.py:91-107).  The oracle computes both
terms separately; the bound is per element, 1e-10 relative with a floor of
16 ulps of the row's largest entry (parity.check).
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc
from parity import check

from test_gpu_parity import CASES, gpu_opts, ref_opts

torch = pytest.importorskip("torch")
from mpart_amd.torch import kl_loss, kl_objective  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _oracle_kl_input_grad(r, x):
    t = r.Evaluate(x)
    return r.Gradient(x, t) - r.LogDeterminantInputGrad(x)


def _close(got, ref, rel=1e-10):
    """Elementwise: |got - ref| <= rel |ref| + 16 eps max|ref_row| (parity.check)."""
    check(np.asarray(got), np.asarray(ref), rtol=rel, what="input gradient")


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_component_matches_oracle(ci):
    _case("kl_input_grad_component", case=CASES[ci])
    s = Stream(1100 + ci)
    g = mt.CreateTotalOrderComponent(3, 3, gpu_opts(**CASES[ci]))
    c = 0.2 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    x = s.points(3, 600)
    r = orc.CreateComponent(ref_opts(**CASES[ci]), 3, 3)
    r.SetCoeffs(c)
    _close(g.KLInputGrad(x), _oracle_kl_input_grad(r, x))


@pytest.mark.parametrize("in_dim,out_dim,order,case", [(4, 4, 3, {}), (5, 3, 2, dict(posFuncType="Exp")),
                                                       (3, 3, 4, dict(basisType="HermiteFunctions", nugget=0.02)),
                                                       (4, 2, 2, dict(basisLB=-1.0, basisUB=1.0))])
def test_triangular_matches_oracle(in_dim, out_dim, order, case):
    _case("kl_input_grad_triangular", in_dim=in_dim, out_dim=out_dim, order=order, case=case)
    s = Stream(1200 + in_dim * 10 + out_dim)
    g = mt.CreateTriangular(in_dim, out_dim, order, gpu_opts(**case))
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    x = s.points(in_dim, 500)
    r = orc.CreateTriangular(in_dim, out_dim, order, ref_opts(**case))
    r.SetCoeffs(c)
    ref = _oracle_kl_input_grad(r, x)
    spec = g.KLInputGrad(x)
    _close(spec, ref)
    # the scale multiplies every entry, and a strided view of the points is read in place
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda().T
    _close(g.KLInputGrad(xd, scale=0.25).cpu().numpy(), 0.25 * ref)
    # the generic kernel (term records in memory, not specialized code)
    g.SetCodegen(False)
    gen = g.KLInputGrad(x)
    _close(gen, ref)
    _close(gen, spec, rel=1e-13)


def test_large_table_in_global_memory():
    # A 6-D order-6 map: the per-lane value+derivative tables exceed LDS, so
    # the launch keeps them in global memory.
    _case("kl_input_grad_global_tables")
    s = Stream(1300)
    g = mt.CreateTriangular(6, 6, 6, mt.MapOptions())
    c = 0.02 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    x = s.points(6, 300)
    r = orc.CreateTriangular(6, 6, 6, ref_opts())
    r.SetCoeffs(c)
    _close(g.KLInputGrad(x), _oracle_kl_input_grad(r, x))


def test_empty_and_discrete_derivative():
    _case("kl_input_grad_edges")
    g = mt.CreateTriangular(3, 3, 2, mt.MapOptions())
    g.SetCoeffs(0.1 * Stream(1400).normal(g.numCoeffs))
    out = g.KLInputGrad(torch.zeros((3, 0), dtype=torch.float64, device="cuda"))
    assert out.shape == (3, 0)
    d = mt.CreateTriangular(3, 3, 2, mt.MapOptions(contDeriv=False))
    d.SetCoeffs(0.1 * Stream(1400).normal(d.numCoeffs))
    with pytest.raises(ValueError, match="Discrete derivative"):
        d.KLInputGrad(Stream(1401).points(3, 10))
    with pytest.raises(ValueError):
        g.KLInputGrad(Stream(1402).points(2, 10))


def test_torch_objective_input_gradient():
    # KLObjective's point gradient against the two-pass autograd path
    # (Gradient + LogDeterminantInputGrad through MpartTorchAutograd) and
    # against central differences of the objective in single coordinates.
    _case("kl_input_grad_torch")
    tmap = mt.CreateTriangular(4, 4, 3, mt.MapOptions(quadAbsTol=1e-12, quadRelTol=1e-12))
    s = Stream(1500)
    c = (0.1 * torch.from_numpy(s.normal(tmap.numCoeffs))).cuda().requires_grad_(True)
    x0 = torch.from_numpy(s.points(4, 800)).cuda()
    x = x0.clone().requires_grad_(True)
    loss = kl_objective(tmap, x, c)
    loss.backward()
    gx, gc = x.grad.detach().clone(), c.grad.detach().clone()
    x2 = x0.clone().requires_grad_(True)
    c.grad = None
    loss2 = kl_loss(tmap, x2, c)
    loss2.backward()
    assert abs(loss.item() - loss2.item()) <= 1e-12 * abs(loss2.item())
    assert (gx - x2.grad).abs().max().item() <= 1e-10 * x2.grad.abs().max().item()
    assert (gc - c.grad).abs().max().item() <= 1e-10 * c.grad.abs().max().item()
    h = 1e-5
    with torch.no_grad():
        for d, n in [(0, 0), (1, 17), (2, 400), (3, 799), (3, 5)]:
            xp = x0.clone()
            xm = x0.clone()
            xp[d, n] += h
            xm[d, n] -= h
            fd = (kl_objective(tmap, xp, c).item() - kl_objective(tmap, xm, c).item()) / (2 * h)
            # the objective is a mean over 800 points: entries are O(1e-3), FD error O(1e-10)
            assert abs(fd - gx[d, n].item()) <= 1e-6 * abs(gx[d, n].item()) + 1e-9, (d, n, fd)
    # without requires_grad on the points no point gradient is formed
    c.grad = None
    kl_objective(tmap, x0, c).backward()
    assert (c.grad - gc).abs().max().item() == 0.0


@pytest.mark.parametrize("ci", [0, 2, 6, 9, 11, 13])
def test_input_vjp_matches_oracle(ci):
    # sens * grad T + lsens * grad logdet per point, with either term absent;
    # specialized and generic kernels.
    _case("input_vjp", case=CASES[ci])
    s = Stream(1600 + ci)
    g = mt.CreateTriangular(4, 3, 3, gpu_opts(**CASES[ci]))
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    x = s.points(4, 400)
    sens = s.normal(3 * 400).reshape(3, 400)
    ls = s.normal(400)
    r = orc.CreateTriangular(4, 3, 3, ref_opts(**CASES[ci]))
    r.SetCoeffs(c)
    gr = r.Gradient(x, sens)
    lg = r.LogDeterminantInputGrad(x)
    for codegen in (True, False):
        g.SetCodegen(codegen)
        _close(g.InputVJP(x, sens, ls), gr + ls[None, :] * lg)
        _close(g.InputVJP(x, sens, None), gr)
        _close(g.InputVJP(x, None, ls), ls[None, :] * lg)
        # Gradient / LogDeterminantInputGrad (specialized: the same fused kernel)
        _close(g.Gradient(x, sens), gr)
        _close(g.LogDeterminantInputGrad(x), lg)


def test_torch_backward_input_gradient_is_one_fused_call():
    # MpartTorchAutograd's point gradient (InputVJP) against the oracle's
    # Gradient + lsens * LogDeterminantInputGrad for the same sensitivities.
    _case("input_vjp_torch")
    from mpart_amd.torch import MpartTorchAutograd
    s = Stream(1700)
    tmap = mt.CreateTriangular(3, 3, 3, mt.MapOptions())
    c = 0.2 * s.normal(tmap.numCoeffs)
    tmap.SetCoeffs(c)
    xn = s.points(3, 300)
    x = torch.from_numpy(xn).cuda().requires_grad_(True)
    ct = torch.from_numpy(c).cuda()
    y, ld = MpartTorchAutograd.apply(x, ct, tmap, True)
    w = torch.from_numpy(s.normal(3 * 300).reshape(3, 300)).cuda()
    v = torch.from_numpy(s.normal(300)).cuda()
    ((y * w).sum() + (ld * v).sum()).backward()
    r = orc.CreateTriangular(3, 3, 3, ref_opts())
    r.SetCoeffs(c)
    ref = r.Gradient(xn, w.cpu().numpy()) + v.cpu().numpy()[None, :] * r.LogDeterminantInputGrad(xn)
    _close(x.grad.cpu().numpy(), ref)


@pytest.mark.parametrize("ci", [0, 1, 2, 6, 10, 12])
def test_one_pass_value_and_both_gradients(ci):
    # KLValueAndGradsImpl / VJPImpl (one specialized kernel, or the generic
    # fused kernel followed by the input VJP) against the separate entry points.
    _case("kl_value_and_grads", case=CASES[ci])
    s = Stream(1800 + ci)
    g = mt.CreateTriangular(5, 4, 3, gpu_opts(**CASES[ci]))
    g.SetCoeffs(0.1 * s.normal(g.numCoeffs))
    n = 3000
    x = torch.from_numpy(s.points(5, n)).cuda()
    sens = torch.from_numpy(s.normal(4 * n).reshape(4, n)).cuda()
    ls = torch.from_numpy(s.normal(n)).cuda()
    for codegen in (True, False):
        g.SetCodegen(codegen)
        loss0, grad0 = g.KLValueAndGrad(x)
        gx0 = g.KLInputGrad(x, scale=1.0 / n)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        grad = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
        gx = torch.full((5, n), 7.0, dtype=torch.float64, device="cuda")
        g.KLValueAndGradsImpl(x, loss, grad, gx)
        _close(loss.cpu().numpy(), loss0.cpu().numpy(), rel=1e-13)
        _close(grad.cpu().numpy(), grad0.cpu().numpy(), rel=1e-13)
        _close(gx.cpu().numpy(), gx0.cpu().numpy(), rel=1e-13)
        # without the input output it is the plain fused pass
        g.KLValueAndGradsImpl(x, loss, grad, None)
        _close(grad.cpu().numpy(), grad0.cpu().numpy(), rel=1e-13)
        # VJP: both sides of the torch backward
        for sv, lv in ((sens, ls), (sens, None), (None, ls)):
            cg0 = g.CoeffVJP(x, sv, lv)
            gi0 = g.InputVJP(x, sv, lv)
            cg = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
            gi = torch.full((5, n), 7.0, dtype=torch.float64, device="cuda")
            g.VJPImpl(x, sv, lv, cg, gi)
            _close(cg.cpu().numpy(), cg0.cpu().numpy(), rel=1e-13)
            _close(gi.cpu().numpy(), gi0.cpu().numpy(), rel=1e-13)


def test_one_pass_large_batch_and_graph_free_reuse():
    # Enough points for a persistent grid that loops over many tiles, a
    # point-major strided view, and a second call reusing the buffers.
    _case("kl_value_and_grads_large")
    s = Stream(1900)
    g = mt.CreateTriangular(8, 8, 3, mt.MapOptions())
    g.SetCoeffs(0.1 * s.normal(g.numCoeffs))
    n = 300_000
    xpm = torch.from_numpy(np.ascontiguousarray(s.points(8, n).T)).cuda()
    x = xpm.T
    gx0 = g.KLInputGrad(x, scale=1.0 / n)
    loss0, grad0 = g.KLValueAndGrad(x)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(g.numCoeffs, dtype=torch.float64, device="cuda")
    gxt = torch.zeros((n, 8), dtype=torch.float64, device="cuda")
    for _ in range(2):
        g.KLValueAndGradsImpl(x, loss, grad, gxt.T)
        _close(gxt.T.cpu().numpy(), gx0.cpu().numpy(), rel=1e-13)
        _close(grad.cpu().numpy(), grad0.cpu().numpy(), rel=1e-13)
        _close(loss.cpu().numpy(), loss0.cpu().numpy(), rel=1e-13)
