"""GPU parity for map shapes at and past the unrolled fast paths.

Terms with more off-diagonal factors than the unrolled term-record loops
handle (eight) take the run-time loop; the reference has no such limit
(MParT/MultivariateExpansionWorker.h:375-397 multiplies every nonzero factor
of a term).  Each case runs the generic kernels and the specialized ones
against the oracle: values and gradients to 1e-10 relative per element,
identical adaptive trees, Inverse to 1e-9 absolute.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc

from test_gpu_parity import close, gpu_opts, ref_opts

pytestmark = pytest.mark.gpu


def _csr(rows):
    """Compressed (nzStarts, nzDims, nzOrders) of dense multi-indices."""
    starts, dims, orders = [0], [], []
    for r in rows:
        for d, o in enumerate(r):
            if o:
                dims.append(d)
                orders.append(o)
        starts.append(len(dims))
    return (np.array(starts, np.uint32), np.array(dims, np.uint32), np.array(orders, np.uint32))


def _many_factor_rows(dim):
    """Constant, every linear term, x_k^2, and terms with dim-1 and dim-2
    off-diagonal factors (with and without the last coordinate)."""
    rows = [[0] * dim]
    for d in range(dim):
        rows.append([int(i == d) for i in range(dim)])
    rows.append([0] * (dim - 1) + [2])
    rows.append([1] * dim)                       # dim-1 off-diagonal factors and x_k
    rows.append([1] * (dim - 1) + [0])           # dim-1 off-diagonal factors only
    rows.append([2] + [1] * (dim - 3) + [0, 1])  # dim-2 factors, one of order 2
    return rows


@pytest.mark.parametrize("dim", [10, 12])
@pytest.mark.parametrize("codegen", [False, True])
def test_terms_with_many_factors(dim, codegen):
    print(f"[gpu-case] many_factors dim={dim} codegen={codegen}", flush=True)
    s, d, o = _csr(_many_factor_rows(dim))
    mset = mt.FixedMultiIndexSet.from_csr(dim, s, d, o)
    g = mt.CreateComponent(mset, gpu_opts())
    g.SetCodegen(codegen)
    r = orc.CreateComponent(ref_opts(), dim, starts=s, dims=d, orders=o)
    st = Stream(1500 + dim)
    x = 0.7 * st.points(dim, 700)
    c = 0.3 * st.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    yg, neg = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    assert np.array_equal(neg, ner), f"tree mismatch at {np.count_nonzero(neg != ner)} points"
    close(yg, yr, what="Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")
    sens = st.normal(x.shape[1])[None, :]
    close(g.CoeffGrad(x, sens), r.CoeffGrad(x, sens), what="CoeffGrad")
    close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad")
    close(g.Gradient(x, sens), r.Gradient(x, sens), what="Gradient")
    close(g.LogDeterminantInputGrad(x), r.LogDeterminantInputGrad(x), what="LogDeterminantInputGrad")
    lr, gr = r.KL(x)
    lf, gf = g.KLValueAndGrad(x)
    close(lf, lr, what="KL loss")
    close(gf, gr, what="KL grad")
    tgt = st.normal(x.shape[1])[None, :]
    xi_g = g.Inverse(x, tgt)
    xi_r = r.Inverse(x, tgt)
    fin = np.isfinite(xi_r)
    assert np.array_equal(fin, np.isfinite(xi_g))
    assert np.max(np.abs(xi_g[fin] - xi_r[fin])) <= 1e-9


# Quadrature rules of more than 17 points (quadPts above 17 for the fixed
# rule; the nested adaptive rule has 2^(L+1) + 1 fine points with
# L = floor(log2(quadPts - 2)), MParT/MapFactoryImpl4.cpp:18): the rule is read
# from device memory and the interior node values of an accepted interval are
# evaluated again for the adjoint.
BIG_RULES = [
    dict(quadType="AdaptiveClenshawCurtis", quadPts=20, quadAbsTol=1e-9, quadRelTol=1e-9),  # nc 17, nf 33
    dict(quadType="AdaptiveClenshawCurtis", quadPts=40, quadAbsTol=1e-9, quadRelTol=1e-9,
         basisType="HermiteFunctions", posFuncType="Exp"),                                  # nc 33, nf 65
    dict(quadType="ClenshawCurtis", quadPts=25),
    dict(quadType="ClenshawCurtis", quadPts=48, basisType="PhysicistHermite", posFuncType="Exp"),
]


@pytest.mark.parametrize("ci", range(len(BIG_RULES)))
@pytest.mark.parametrize("codegen", [False, True])
def test_quadrature_rules_above_17_points(ci, codegen):
    case = BIG_RULES[ci]
    print(f"[gpu-case] big_rule case={case} codegen={codegen}", flush=True)
    st = Stream(1600 + ci)
    g = mt.CreateTotalOrderComponent(3, 3, gpu_opts(**case))
    g.SetCodegen(codegen)
    r = orc.CreateComponent(ref_opts(**case), 3, 3)
    x = st.points(3, 600)
    c = 0.2 * st.normal(g.numCoeffs)
    g.SetCoeffs(c)
    r.SetCoeffs(c)
    yg, neg = g.EvaluateCounted(x)
    yr, ner = r.Evaluate(x, return_nevals=True)
    assert np.array_equal(neg, ner), f"tree mismatch at {np.count_nonzero(neg != ner)} points"
    close(yg, yr, what="Evaluate")
    close(g.LogDeterminant(x), r.LogDeterminant(x), what="LogDeterminant")
    sens = st.normal(x.shape[1])[None, :]
    close(g.CoeffGrad(x, sens), r.CoeffGrad(x, sens), what="CoeffGrad")
    close(g.LogDeterminantCoeffGrad(x), r.LogDeterminantCoeffGrad(x), what="LogDeterminantCoeffGrad")
    close(g.Gradient(x, sens), r.Gradient(x, sens), what="Gradient")
    close(g.LogDeterminantInputGrad(x), r.LogDeterminantInputGrad(x), what="LogDeterminantInputGrad")
    tgt = st.normal(x.shape[1])[None, :]
    xi_g = g.Inverse(x, tgt)
    xi_r = r.Inverse(x, tgt)
    fin = np.isfinite(xi_r)
    assert np.array_equal(fin, np.isfinite(xi_g))
    assert np.max(np.abs(xi_g[fin] - xi_r[fin])) <= 1e-9
    # a triangular map: the fused KL pass and its plain structure
    gt = mt.CreateTriangular(3, 3, 2, gpu_opts(**case))
    gt.SetCodegen(codegen)
    rt = orc.CreateTriangular(3, 3, 2, ref_opts(**case))
    ct = 0.1 * st.normal(gt.numCoeffs)
    gt.SetCoeffs(ct)
    rt.SetCoeffs(ct)
    lr, gr = rt.KL(x)
    lf, gf = gt.KLValueAndGrad(x)
    lp, gp = gt.KLValueAndGrad(x, plain=True)
    close(lf, lr, what="KL loss fused")
    close(gf, gr, what="KL grad fused")
    close(lp, lr, what="KL loss plain")
    close(gp, gr, what="KL grad plain")


def test_quadrature_rule_above_17_points_batched():
    print("[gpu-case] big_rule batched", flush=True)
    torch = pytest.importorskip("torch")
    case = BIG_RULES[0]
    st = Stream(1700)
    maps, refs, xs = [], [], []
    for k, (dim, order) in enumerate([(2, 2), (3, 3), (4, 2)]):
        g = mt.CreateTriangular(dim, dim, order, gpu_opts(**case))
        r = orc.CreateTriangular(dim, dim, order, ref_opts(**case))
        c = 0.1 * st.normal(g.numCoeffs)
        g.SetCoeffs(c)
        r.SetCoeffs(c)
        maps.append(g)
        refs.append(r)
        xs.append(torch.from_numpy(st.points(dim, 300 + 100 * k)).cuda())
    b = mt.MapBatch(maps)
    ys, lds = b.Evaluate(xs, logdet=True)
    for y, ld, r, x in zip(ys, lds, refs, xs):
        xn = x.cpu().numpy()
        close(y.cpu().numpy(), r.Evaluate(xn), what="batched Evaluate")
        close(ld.cpu().numpy(), r.LogDeterminant(xn), what="batched LogDeterminant")
    for sched in (1, 2):
        b.SetInverseSchedule(sched)
        outs = b.Inverse([torch.zeros((0, x.shape[1]), dtype=torch.float64, device="cuda") for x in xs], ys)
        for o, r, y in zip(outs, refs, ys):
            xr = r.Inverse(np.zeros((0, y.shape[1])), y.cpu().numpy())
            xg = o.cpu().numpy()
            fin = np.isfinite(xr)
            assert np.array_equal(fin, np.isfinite(xg)), sched
            assert np.max(np.abs(xg[fin] - xr[fin])) <= 1e-9, sched
