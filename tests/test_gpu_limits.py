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
