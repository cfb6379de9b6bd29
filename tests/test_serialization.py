"""The reference's on-disk format (cereal binary archives of MapOptions,
FixedMultiIndexSet and map coefficients), host side.

Parity unpinned: the reference's tests write their archives to temporary
files (bindings/python/tests/test_Serialization.py), so no reference-written
archive exists to compare with, and cereal is not installed here.  These
tests fix the layout that cereal's BinaryOutputArchive gives the reference's
save functions (raw little-endian values, one-byte bools, int-sized enums,
uint64 string lengths, views as label + unsigned length + elements) and the
round trips; the GPU test rebuilds a component from the three files as the
reference's test does.
"""
import math
import struct

import numpy as np

import mpart_amd as mt
from mpart_amd import serialization as ser


def test_map_options_archive_layout_and_round_trip(tmp_path):
    o = mt.MapOptions(basisType=mt.BasisTypes.HermiteFunctions, basisLB=-2.5, posFuncType=mt.PosFuncTypes.Exp,
                      quadType=mt.QuadTypes.AdaptiveClenshawCurtis, quadAbsTol=1e-9, quadRelTol=2e-9,
                      quadMaxSub=12, quadMinSub=1, quadPts=9, contDeriv=False, basisNorm=False, nugget=0.25)
    b = ser.options_to_bytes(o)
    assert len(b) == 4 + 8 + 8 + 4 + 4 + 8 + 8 + 4 + 4 + 4 + 1 + 1 + 8
    assert struct.unpack_from("<i", b, 0)[0] == 2
    assert struct.unpack_from("<d", b, 4)[0] == -2.5 and struct.unpack_from("<d", b, 12)[0] == math.inf
    assert struct.unpack_from("<iidd", b, 20) == (0, 2, 1e-9, 2e-9)
    assert struct.unpack_from("<III??d", b, 44) == (12, 1, 9, False, False, 0.25)
    f = tmp_path / "opts.mt"
    o.Serialize(str(f))
    o2 = mt.MapOptions().Deserialize(str(f))
    for k in mt.MapOptions._FIELDS:
        assert getattr(o2, k) == getattr(o, k), k
    assert isinstance(o2.basisType, mt.BasisTypes)


def test_multiset_archive_layout_and_round_trip(tmp_path):
    s = mt.FixedMultiIndexSet(3, 2)
    b = ser.multiset_to_bytes(s)
    assert struct.unpack_from("<I?", b, 0) == (3, True)
    n = struct.unpack_from("<Q", b, 5)[0]
    assert b[13:13 + n] == b"nzStarts"
    assert struct.unpack_from("<I", b, 13 + n)[0] == s.Size() + 1
    f = tmp_path / "fmset.mt"
    s.Serialize(str(f))
    s2 = mt.FixedMultiIndexSet().Deserialize(str(f))
    assert s2.dim == 3
    for a in ("nzStarts", "nzDims", "nzOrders"):
        assert np.array_equal(getattr(s2, a), getattr(s, a))
    assert np.array_equal(s2.MaxDegrees(), s.MaxDegrees())


def test_map_coefficient_archive(tmp_path):
    c = np.linspace(-1.0, 2.0, 7)
    b = ser.map_to_bytes(5, 2, c)
    assert struct.unpack_from("<III", b, 0) == (5, 2, 7)
    f = tmp_path / "comp.mt"
    f.write_bytes(b)
    i, o, c2 = mt.DeserializeMap(str(f))
    assert (i, o) == (5, 2) and np.array_equal(c2, c)


def test_map_byte_string_layout_and_round_trip():
    # ToBytes' layout: magic, version, triangular flag, the MapOptions archive,
    # the component count, each FixedMultiIndexSet archive with its length,
    # then the map archive (host side; the GPU test round-trips real maps).
    o = mt.MapOptions(basisType=mt.BasisTypes.PhysicistHermite, quadType=mt.QuadTypes.AdaptiveClenshawCurtis,
                      quadPts=9, nugget=0.5, basisLB=-2.0, basisUB=3.0)
    sets = [mt.FixedMultiIndexSet(1, 2), mt.FixedMultiIndexSet(2, 3)]
    c = np.linspace(0.0, 1.0, sets[0].Size() + sets[1].Size())
    b = ser.map_state_to_bytes(o, True, sets, 2, 2, c)
    assert b[:8] == b"MPARTAMD" and struct.unpack_from("<IB", b, 8) == (1, 1)
    assert b[13:13 + 66] == ser.options_to_bytes(o)
    assert struct.unpack_from("<I", b, 79)[0] == 2
    n0 = struct.unpack_from("<Q", b, 83)[0]
    assert b[91:91 + n0] == ser.multiset_to_bytes(sets[0])
    assert b.endswith(ser.map_to_bytes(2, 2, c))
    o2 = mt.MapOptions()
    tri, got, i, j, c2 = ser.map_state_from_bytes(b, o2)
    assert tri and (i, j) == (2, 2) and np.array_equal(c2, c)
    for s, (dim, st, dm, od) in zip(sets, got):
        assert dim == s.dim and np.array_equal(st, s.nzStarts) and np.array_equal(dm, s.nzDims)
        assert np.array_equal(od, s.nzOrders)
    for k in mt.MapOptions._FIELDS:
        assert getattr(o2, k) == getattr(o, k), k
    try:
        ser.map_state_from_bytes(b"\x00" * 40, o2)
    except ValueError:
        pass
    else:
        raise AssertionError("a foreign byte string must be refused")
