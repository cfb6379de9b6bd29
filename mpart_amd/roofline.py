"""Algorithmic work model of the fused forward+adjoint pass (for roofline figures).

Counting convention: add/mul = 1, FMA = 2, div/sqrt/exp/log = 1 each,
compare/select = 0.  Unit: one point of one map, all components.

Per component c with last-coordinate degree p, evaluation count E_c (taken
from the adaptive trees of the actual inputs) and, for adaptive Simpson,
V_c = (E_c - 3) / 2 visited intervals of which L_c = (V_c + 1) / 2 are accepted:

  node      = 3 max(p-2, 0) + 3p + g + 3          basis recurrence, d_k f, bijector,
                                                  F = x (g + nugget), h = x g'
            with g = 8 (SoftPlus: exp, |.|, add, log, max, add, div, add) or 1 (Exp)
  interval  = 20                                  coarse/fine Simpson sums, error test
  leaf      = 5 (3 max(p-2, 0) + 3p + 2)          adjoint: B'_q at 5 nodes, S'_q accumulation
  fixed     = sum_j (f_j + 1)                     C_q = sum w_j P_j   (f_j off-diagonal factors)
            + sum_j (f_j + 1)                     gradient: P_j G_q and the point reduction
            + 2 (p+1) + 3p + 4 + 6p               f(x_<k, 0), log-det, G_q
plus the basis table, 3 flops per (coordinate, order) entry.

Bytes per point: 8 D (points read once); outputs are numCoeffs + 1 doubles per
launch (reduced), so HBM traffic is ~8 D bytes per point.
"""
import numpy as np


def _node(p, exp_bijector):
    return 3 * max(p - 2, 0) + 3 * p + (1 if exp_bijector else 8) + 3


def _leaf(p):
    return 5 * (3 * max(p - 2, 0) + 3 * p + 2)


def component_fixed(mset, p):
    f = np.diff(mset.nzStarts).astype(np.int64)
    # off-diagonal factors of each term = nonzeros minus a last-coordinate entry
    last_dim = mset.dim - 1
    has_last = np.zeros(len(f), dtype=np.int64)
    for t in range(len(f)):
        a, b = mset.nzStarts[t], mset.nzStarts[t + 1]
        if b > a and mset.nzDims[b - 1] == last_dim:
            has_last[t] = 1
    off = f - has_last
    return int(2 * np.sum(off + 1) + 2 * (p + 1) + 3 * p + 4 + 6 * p)


def fused_flops(msets, counts, exp_bijector=False, quad="AdaptiveSimpson"):
    """Total algorithmic flops of one fused pass.

    msets: list of FixedMultiIndexSet, one per component.
    counts: (ncomps, N) integrand-evaluation counts of the inputs.
    """
    total = 0.0
    D = msets[-1].dim
    N = counts.shape[1]
    table = 0
    for d in range(D - 1):
        table = max(table, 0)
    maxdeg = np.zeros(D, dtype=np.int64)
    for m in msets:
        md = m.MaxDegrees()
        for d in range(m.dim - 1):
            maxdeg[d] = max(maxdeg[d], md[d])
    total += 3.0 * float(maxdeg.sum()) * N
    for c, m in enumerate(msets):
        p = int(m.MaxDegrees()[m.dim - 1])
        E = counts[c].astype(np.float64)
        if quad == "AdaptiveSimpson":
            V = (E - 3.0) / 2.0
            L = (V + 1.0) / 2.0
        else:
            V = (E - 2.0) / 3.0
            L = (V + 1.0) / 2.0
        total += float(np.sum(E)) * _node(p, exp_bijector)
        total += float(np.sum(V)) * 20.0
        total += float(np.sum(L)) * _leaf(p)
        total += float(component_fixed(m, p)) * N
    return total
