"""Algorithmic work models of the fused forward+adjoint pass (roofline figures).

Two models.  ``algo_flops`` counts the reference's algorithm (the per-node
term contraction of MParT's MonotoneIntegrand / MultivariateExpansionWorker,
one point at a time): it is the figure the roofline's ``achieved`` uses.
``fused_flops`` counts this library's regrouped algorithm (C_q sums, O(p) per
node), which does less arithmetic for the same result; it is reported beside.

algo_flops (reference algorithm).  Unit: one point of one map, all
components; convention add/mul = 1, FMA = 2, div/sqrt/exp/log = 1,
compare = 0.  Per component c with term set J, diagonal terms Diag (a nonzero
last-coordinate order), nnz_j nonzeros of term j, last-coordinate degree p,
E_c integrand evaluations and V_c visited intervals of the point's adaptive
tree (Simpson: E = 3 + 2V; Clenshaw-Curtis 3/5: E = 2 + 3V):

  FLOPs_c   = E_c phi_node + 20 V_c + phi_fixed
  phi_node  = phi_basis(p) + sum_{j in Diag} (nnz_j + 2) + phi_g + 2
              (+ 3 |Diag| + 2 in adjoint mode: scaling and accumulation of
              grad_w d_k f, and g')
  phi_basis = 9 max(p - 1, 0) + 2 (p + 1)     values, derivatives, normalization
  phi_g     = 4 (SoftPlus) or 1 (Exp)
  phi_fixed = sum_{j in J} (nnz_j + 2) + |Diag|   f(x_<k, 0), grad_w f(x_<k, 0),
                                                the log-det and its gradient

(the same phi_basis is used for every basis family).  On the 8-D order-3 map
at five evaluations per component this gives 9 562 flops per point.

fused_flops (regrouped algorithm).  Same convention and unit.
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


def _term_structure(mset):
    """(nnz per term, diagonal-term mask) of a FixedMultiIndexSet."""
    st = np.asarray(mset.nzStarts, dtype=np.int64)
    dims = np.asarray(mset.nzDims, dtype=np.int64)
    nnz = np.diff(st)
    last = mset.dim - 1
    diag = np.array([st[j + 1] > st[j] and dims[st[j + 1] - 1] == last for j in range(len(nnz))], dtype=bool)
    return nnz, diag


def algo_flops(msets, counts, adjoint=True, exp_bijector=False, quad="AdaptiveSimpson"):
    """Total flops of one pass over the points by the reference's algorithm
    (module docstring).  msets: one FixedMultiIndexSet per component;
    counts: (ncomps, N) integrand-evaluation counts of the inputs."""
    total = 0.0
    N = counts.shape[1]
    for c, m in enumerate(msets):
        nnz, diag = _term_structure(m)
        p = int(m.MaxDegrees()[m.dim - 1])
        node = 9 * max(p - 1, 0) + 2 * (p + 1) + float(np.sum(nnz[diag] + 2)) + (1 if exp_bijector else 4) + 2
        if adjoint:
            node += 3 * int(diag.sum()) + 2
        fixed = float(np.sum(nnz + 2)) + int(diag.sum())
        E = counts[c].astype(np.float64)
        V = (E - 3.0) / 2.0 if quad == "AdaptiveSimpson" else (E - 2.0) / 3.0
        total += float(np.sum(E)) * node + 20.0 * float(np.sum(V)) + fixed * N
    return total
