"""CPU-side checks of the C ABI library (no GPU needed): it loads, exports every
entry point declared in include/mpart_amd.h, enumerates multi-index sets
bit-exactly like the oracle (FixedMultiIndexSet.cpp:487-549), and the host
helpers agree with the reference defaults."""
import ctypes
import itertools
import math
import os
import re

import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream, splitmix64, banana_points, student_t_points
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "mpart_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(mpart_[a-z_0-9]+)\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mt.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # The Python front end binds every declared entry point.
    assert set(syms) <= set(mt.C_EXPORTS)


@pytest.mark.parametrize("dim,order", [(1, 0), (1, 5), (2, 3), (3, 4), (6, 4), (8, 3), (12, 3), (16, 2), (5, 6)])
def test_total_order_enumeration_bit_exact(dim, order):
    s = mt.FixedMultiIndexSet(dim, order)
    starts, dims, orders, maxdeg = orc.total_order(dim, order)
    assert np.array_equal(s.nzStarts, starts)
    assert np.array_equal(s.nzDims, dims)
    assert np.array_equal(s.nzOrders, orders)
    assert np.array_equal(s.MaxDegrees(), maxdeg)
    assert s.Size() == math.comb(dim + order, order)
    if dim <= 6:
        multis = [a for a in itertools.product(range(order + 1), repeat=dim) if sum(a) <= order]
        assert [tuple(s.IndexToMulti(i)) for i in range(s.Size())] == multis


def test_config_term_counts():
    # C4 (literal reading): component k has order max(1, 6 - k // 6) and k + 1 inputs.
    counts = [mt.FixedMultiIndexSet(k + 1, max(1, 6 - k // 6)).Size() for k in range(32)]
    assert counts[0] == 7 and sum(counts) == 62572 and max(counts) == 7315
    # C3: 494 coefficients.
    assert sum(mt.FixedMultiIndexSet(d, 3).Size() for d in range(1, 9)) == 494


def test_map_options_defaults():
    # MParT/MapOptions.h:61-123
    o = mt.MapOptions()
    assert o.basisType == mt.BasisTypes.ProbabilistHermite
    assert o.posFuncType == mt.PosFuncTypes.SoftPlus
    assert o.quadType == mt.QuadTypes.AdaptiveSimpson
    assert (o.quadAbsTol, o.quadRelTol, o.quadMaxSub, o.quadMinSub, o.quadPts) == (1e-6, 1e-6, 30, 0, 5)
    assert o.contDeriv is True and o.basisNorm is True and o.nugget == 0.0
    with pytest.raises(ValueError):
        mt.MapOptions(notAField=1)


def test_splitmix64_known_answers():
    # Published SplitMix64 outputs for seed 0.
    z = splitmix64(0, 0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_stream_fixture():
    # First values of the documented streams (seed 0), kept as a fixture.
    fx = np.load(os.path.join(ROOT, "tests", "golden", "streams.npz"))
    s = Stream(0)
    assert np.array_equal(s.uniform(4), fx["uniform4"])
    assert np.array_equal(Stream(0).normal(4), fx["normal4"])
    assert np.array_equal(banana_points(4)[:, 0], fx["banana_col0"])
    assert np.array_equal(student_t_points(3, 2)[:, 0], fx["student_col0"])


def test_banana_generator_recursion():
    x = banana_points(1000)
    eps = Stream(0).normal(8 * 1000).reshape(1000, 8)
    assert np.array_equal(x[0], eps[:, 0])
    np.testing.assert_array_equal(x[1], 0.5 * x[0] ** 2 + np.tanh(0.0) + 0.3 * eps[:, 1])
    np.testing.assert_array_equal(x[5], 0.5 * x[4] ** 2 + np.tanh(x[3]) + 0.3 * eps[:, 5])
