"""Elementwise comparison of GPU results with the oracle's.

Bound per element:  |gpu - ref| <= rtol * |ref| + atol,
with rtol the relative tolerance the comparison states (1e-10 for values and
gradients) and atol an absolute floor stated per output.  The floor only
matters where |ref| sits at cancellation level: an element that is the
difference of terms of size S carries a rounding error of a few ulps of S
whatever its own size, so atol is a few ulps of the row's scale
(``ulps * eps * max|ref_row|``), far below rtol * |ref| for every element
whose size is within 1e-5 of the row maximum.

Adaptive-quadrature outputs carry their subdivision trees: where the GPU
and the oracle take different trees at a point (a "flip": an interval whose
error estimate sits within rounding of the tolerance), the two values differ
by up to the quadrature tolerance, not by rounding.  Such points are counted
and reported, and checked against the quadrature tolerance instead.
"""
import numpy as np

EPS = np.finfo(np.float64).eps


def stats(gpu, ref, rtol=1e-10, ulps=16.0, what="", mask=None):
    """Elementwise error statistics of gpu against ref (same shape).

    Returns a dict: n (finite elements compared), max_rel (max |d|/|ref| over
    elements with ref != 0), max_rel_above_floor (the same over elements whose
    |ref| exceeds atol / rtol, i.e. where the floor plays no part), atol (per
    row), worst (max |d| / bound), n_floor (elements where the floor is
    larger than rtol * |ref|), n_bad (elements over the bound).  mask (bool,
    broadcastable to the points axis) leaves points out of the comparison
    (tree flips, compared separately)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert gpu.shape == ref.shape, (what, gpu.shape, ref.shape)
    fin = np.isfinite(ref)
    if not np.array_equal(fin, np.isfinite(gpu)):
        bad = int(np.count_nonzero(fin != np.isfinite(gpu)))
        raise AssertionError(f"{what}: non-finite pattern differs at {bad} elements")
    if mask is not None:
        fin = fin & ~np.broadcast_to(mask, ref.shape)
    r = np.where(fin, ref, 0.0)
    d = np.abs(np.where(fin, gpu - ref, 0.0))
    if ref.ndim == 2:
        scale = np.max(np.abs(r), axis=1, keepdims=True)
    else:
        scale = np.max(np.abs(r)) if r.size else 0.0
    atol = ulps * EPS * scale
    bound = rtol * np.abs(r) + atol
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(fin & (r != 0), d / np.abs(r), 0.0)
        above = fin & (rtol * np.abs(r) >= atol)
        worst = float(np.max(np.where(fin, d / np.maximum(bound, 1e-300), 0.0))) if d.size else 0.0
    return {
        "what": what,
        "n": int(np.count_nonzero(fin)),
        "max_rel": float(np.max(rel)) if rel.size else 0.0,
        "max_rel_above_floor": float(np.max(np.where(above, rel, 0.0))) if rel.size else 0.0,
        "atol_max": float(np.max(atol)) if np.size(atol) else 0.0,
        "n_floor": int(np.count_nonzero(fin & ~above)),
        "worst": worst,
        "n_bad": int(np.count_nonzero(d > bound)),
    }


def check(gpu, ref, rtol=1e-10, ulps=16.0, what="", mask=None):
    """stats(), printed, and asserted: every compared element within its bound."""
    s = stats(gpu, ref, rtol, ulps, what, mask)
    print(f"[parity] {what}: n={s['n']} max_rel={s['max_rel']:.3g} "
          f"max_rel(|ref| above floor)={s['max_rel_above_floor']:.3g} floor={ulps:g} ulp of row max "
          f"(<= {s['atol_max']:.3g}, used by {s['n_floor']}) worst/bound={s['worst']:.3g}", flush=True)
    assert s["n_bad"] == 0, f"{what}: {s['n_bad']} elements over the bound (worst {s['worst']:.3g} x bound)"
    return s
