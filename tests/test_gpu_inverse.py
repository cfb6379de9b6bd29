"""Inverse with per-lane solve scheduling (the single-map path) against the
lane-per-point sequential solver (the batched path's schedule 2, which runs
bracket + ITP for one point per lane from start to end) and the oracle; the
batched queue path (schedule 1: the same solver over the points of many maps)
against the oracle.

Both paths evaluate the same nodes in the same accumulation order and take
the same branches, so their roots agree bit for bit; against the oracle the
bound is 1e-9 absolute.  Targets are drawn wide (N(0, 3^2)) so that some lie
outside the component's range: those solves walk far out in the bracket
search and end in NaN, which gives the lanes of a wavefront very different
amounts of work.
"""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream
from oracle import oracle as orc

from test_gpu_parity import CASES, gpu_opts, ref_opts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[gpu-case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _component(case, seed, n, dim=3, order=3, scale=0.2, nan_at=()):
    s = Stream(seed)
    g = mt.CreateTotalOrderComponent(dim, order, gpu_opts(**case))
    x = s.points(dim, n)
    c = scale * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    tgt = 3.0 * s.normal(n)[None, :]
    for i, j in nan_at:
        x[i, j] = np.nan
    return g, c, x, tgt


def _sequential(g):
    # the lane-per-point solver: each point solved to the end in its lane
    b = mt.MapBatch([g])
    b.SetInverseSchedule(2)
    return b


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_scheduled_inverse_equals_sequential_and_oracle(ci):
    _case("inverse_scheduled", case=CASES[ci])
    g, c, x, tgt = _component(CASES[ci], 900 + ci, 700, nan_at=[(0, 5), (2, 17), (1, 650)])
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    single = g.Inverse(xd, td)
    seq = _sequential(g).Inverse([xd], [td])[0]
    assert torch.equal(torch.isnan(single), torch.isnan(seq))
    assert torch.equal(torch.nan_to_num(single), torch.nan_to_num(seq))
    r = orc.CreateComponent(ref_opts(**CASES[ci]), 3, 3)
    r.SetCoeffs(c)
    ref = r.Inverse(x, tgt)
    got = single.cpu().numpy()
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.isnan(got[0, [5, 17, 650]]).all()
    assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-9
    # the batched queue (per-lane p and outputs): the oracle's roots to 1e-9
    bqb = mt.MapBatch([g])
    bqb.SetInverseSchedule(1)
    bq = bqb.Inverse([xd], [td])[0].cpu().numpy()
    assert np.array_equal(fin, np.isfinite(bq))
    assert np.max(np.abs(bq[fin] - ref[fin])) <= 1e-9


def test_scheduled_inverse_one_column_x1():
    # x1 with one column: every target starts from the same point.
    _case("inverse_broadcast")
    g, c, x, tgt = _component({}, 950, 300)
    x1 = x[:, :1].copy()
    got = g.Inverse(x1, tgt)
    r = orc.CreateComponent(ref_opts(), 3, 3)
    r.SetCoeffs(c)
    ref = r.Inverse(x1, tgt)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-9


@pytest.mark.parametrize("ci", [0, 4])
def test_scheduled_inverse_many_refills(ci):
    # Enough points that every wavefront takes several ranges of point indices
    # from the global counter; the sequential solver is the reference.
    _case("inverse_refills", case=CASES[ci])
    g, c, x, tgt = _component(CASES[ci], 960 + ci, 400_000, dim=4, order=2, scale=0.1)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    out = torch.full_like(td, -7.0)
    g.InverseImpl(xd, td, out)
    seq = _sequential(g).Inverse([xd], [td])[0]
    assert not bool((out == -7.0).any()), "a point was not solved"
    assert torch.equal(torch.isnan(out), torch.isnan(seq))
    assert torch.equal(torch.nan_to_num(out), torch.nan_to_num(seq))
    # the output may alias the initial-guess row of x1
    xa = xd.clone()
    g.InverseImpl(xa, td, xa[-1:])
    assert torch.equal(torch.nan_to_num(xa[-1:]), torch.nan_to_num(seq))


def test_scheduled_triangular_inverse_equals_sequential():
    _case("inverse_triangular_scheduled")
    s = Stream(970)
    g = mt.CreateTriangular(4, 4, 2, mt.MapOptions())
    g.SetCoeffs(0.1 * s.normal(g.numCoeffs))
    n = 5000
    y = torch.from_numpy(2.0 * s.normal(4 * n).reshape(4, n)).cuda()
    empty = torch.zeros((0, n), dtype=torch.float64, device="cuda")
    single = g.Inverse(empty, y)
    seq = _sequential(g).Inverse([empty], [y])[0]
    assert torch.equal(torch.isnan(single), torch.isnan(seq))
    assert torch.equal(torch.nan_to_num(single), torch.nan_to_num(seq))


@pytest.mark.parametrize("ci", [0, 1, 4, 10, 12])
def test_large_scheduled_inverse_equals_sequential(ci):
    # 30 000 points with wide targets over several option cases: every solve
    # is bitwise the sequential one, NaN inputs give NaN.
    _case("inverse_large", case=CASES[ci])
    g, c, x, tgt = _component(CASES[ci], 980 + ci, 30_000, nan_at=[(0, 7), (2, 29_999)])
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tgt).cuda()
    out = torch.full_like(td, -7.0)
    g.InverseImpl(xd, td, out)
    seq = _sequential(g).Inverse([xd], [td])[0]
    assert not bool((out == -7.0).any()), "a point was not solved"
    assert torch.equal(torch.isnan(out), torch.isnan(seq))
    assert torch.equal(torch.nan_to_num(out), torch.nan_to_num(seq))
    assert bool(torch.isnan(out[0, [7, 29_999]]).all())


def test_large_triangular_inverse_equals_sequential():
    _case("inverse_triangular_large")
    s = Stream(990)
    g = mt.CreateTriangular(3, 3, 3, mt.MapOptions())
    g.SetCoeffs(0.1 * s.normal(g.numCoeffs))
    n = 40_000
    y = torch.from_numpy(2.0 * s.normal(3 * n).reshape(3, n)).cuda()
    empty = torch.zeros((0, n), dtype=torch.float64, device="cuda")
    single = g.Inverse(empty, y)
    seq = _sequential(g).Inverse([empty], [y])[0]
    assert torch.equal(torch.isnan(single), torch.isnan(seq))
    assert torch.equal(torch.nan_to_num(single), torch.nan_to_num(seq))


def test_many_term_map_setup_single_equals_batched_and_oracle():
    # A map whose basis table does not fit 256 lanes (9-D, order 4: 37 table
    # slots) takes 4 lanes per point in the Inverse setup, in the single-map
    # path as in the batch: the records, and so the roots, are the same bit
    # for bit.  Against the oracle: 1e-9, except where a comparison of T with
    # the target is decided by its last bits (the 4-lane sums round
    # differently) and the two solvers stop at different points of the
    # tolerance window, which later components of the triangular map inherit
    # (the bound of test_gpu_many_maps_inverse.py: at most 10 % of the
    # coordinates, each within 1e-5).
    _case("inverse_many_terms_setup")
    s = Stream(995)
    g = mt.CreateTriangular(9, 9, 4, mt.MapOptions())
    c = 0.1 * s.normal(g.numCoeffs)
    g.SetCoeffs(c)
    n = 3000
    y = 2.0 * s.normal(9 * n).reshape(9, n)
    yd = torch.from_numpy(y).cuda()
    empty = torch.zeros((0, n), dtype=torch.float64, device="cuda")
    single = g.Inverse(empty, yd)
    b = mt.MapBatch([g])
    b.SetInverseSchedule(1)
    batched = b.Inverse([empty], [yd])[0]
    assert torch.equal(torch.isnan(single), torch.isnan(batched))
    assert torch.equal(torch.nan_to_num(single), torch.nan_to_num(batched))
    r = orc.CreateTriangular(9, 9, 4, ref_opts())
    r.SetCoeffs(c)
    ref = r.Inverse(np.zeros((0, n)), y)
    got = single.cpu().numpy()
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    d = np.abs(got[fin] - ref[fin])
    off = int(np.count_nonzero(d > 1e-9))
    print(f"[gpu-case] roots off the oracle's by more than 1e-9: {off} of {d.size}, max {np.max(d):.3g}", flush=True)
    assert off <= max(1, d.size // 10) and np.max(d) <= 1e-5
