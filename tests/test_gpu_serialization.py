"""Rebuilding a component from the three archives, as the reference's
bindings/python/tests/test_Serialization.py does: MapOptions, the
FixedMultiIndexSet and the coefficients (DeserializeMap); and whole maps
through ToBytes / FromBytes and pickle, as its test_trimap_saveload does
(test_Serialization.py:86-114) (the byte layout is this library's own)."""
import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream

pytestmark = pytest.mark.gpu


def test_component_round_trip_through_archives(tmp_path):
    print("[gpu-case] serialization_round_trip", flush=True)
    opts = mt.MapOptions(basisNorm=False, basisLB=-3.0, basisUB=3.0, quadType=mt.QuadTypes.ClenshawCurtis,
                         quadPts=7)
    mset = mt.FixedMultiIndexSet(2, 3)
    comp = mt.CreateComponent(mset, opts)
    coeffs = np.array([0.5 * (i + 1) for i in range(comp.numCoeffs)]) * 0.1
    comp.SetCoeffs(coeffs)
    mset.Serialize(str(tmp_path / "fmset.mt"))
    opts.Serialize(str(tmp_path / "opts.mt"))
    comp.Serialize(str(tmp_path / "comp.mt"))

    mset2 = mt.FixedMultiIndexSet().Deserialize(str(tmp_path / "fmset.mt"))
    opts2 = mt.MapOptions().Deserialize(str(tmp_path / "opts.mt"))
    inputDim, outputDim, c2 = mt.DeserializeMap(str(tmp_path / "comp.mt"))
    comp2 = mt.CreateComponent(mset2, opts2)
    comp2.SetCoeffs(c2)
    assert (comp2.inputDim, comp2.outputDim) == (inputDim, outputDim) == (2, 1)
    assert np.array_equal(comp2.CoeffMap(), coeffs)
    x = Stream(3).points(2, 500) * 2.0
    assert np.array_equal(comp2.Evaluate(x), comp.Evaluate(x))


def test_map_to_bytes_and_pickle():
    # The reference's test_trimap_saveload: a triangular map of two custom
    # components through ToBytes / TriangularMap.FromBytes and through pickle.
    import pickle

    print("[gpu-case] map_bytes_and_pickle", flush=True)
    opts = mt.MapOptions()
    m1 = mt.CreateComponent(mt.FixedMultiIndexSet.from_csr(1, [0, 0, 1], [0], [1]), opts)          # [0], [1]
    m2 = mt.CreateComponent(mt.FixedMultiIndexSet.from_csr(2, [0, 0, 1, 2], [1, 0], [1, 2]), opts)  # [0,0], [0,1], [2,0]
    tmap = mt.TriangularMap([m1, m2], True)
    tmap.SetCoeffs(Stream(5).normal(tmap.numCoeffs))
    x = Stream(6).points(2, 400)
    for tmap2 in (mt.TriangularMap.FromBytes(tmap.ToBytes()), pickle.loads(pickle.dumps(tmap))):
        assert np.array_equal(tmap2.CoeffMap(), tmap.CoeffMap())
        assert (tmap2.inputDim, tmap2.outputDim, tmap2.NumComponents()) == (2, 2, 2)
        assert np.array_equal(tmap2.Evaluate(x), tmap.Evaluate(x))


@pytest.mark.parametrize("kw", [dict(basisType=mt.BasisTypes.HermiteFunctions, posFuncType=mt.PosFuncTypes.Exp),
                                dict(quadType=mt.QuadTypes.AdaptiveClenshawCurtis, quadPts=9, nugget=0.1),
                                dict(basisLB=-2.0, basisUB=2.5, contDeriv=False, basisNorm=False)])
def test_map_bytes_keep_options(kw):
    import pickle

    print(f"[gpu-case] map_bytes_options {kw}", flush=True)
    g = mt.CreateTriangular(3, 3, 2, mt.MapOptions(**kw))
    g.SetCoeffs(0.1 * Stream(8).normal(g.numCoeffs))
    c = mt.CreateTotalOrderComponent(3, 3, mt.MapOptions(**kw))
    c.SetCoeffs(0.1 * Stream(9).normal(c.numCoeffs))
    x = Stream(10).points(3, 300)
    for m in (g, c):
        m2 = pickle.loads(pickle.dumps(m))
        o1, t1 = m.Options()
        o2, t2 = m2.Options()
        assert t1 == t2 and all(getattr(o1, k) == getattr(o2, k) for k in mt.MapOptions._FIELDS)
        assert np.array_equal(m2.Evaluate(x), m.Evaluate(x))
        assert np.array_equal(m2.LogDeterminant(x), m.LogDeterminant(x))
