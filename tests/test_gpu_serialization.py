"""Rebuilding a component from the three archives, as the reference's
bindings/python/tests/test_Serialization.py does: MapOptions, the
FixedMultiIndexSet and the coefficients (DeserializeMap)."""
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
