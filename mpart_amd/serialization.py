"""The reference's on-disk format for maps: cereal binary archives.

The reference's Python bindings write three files for a component
(bindings/python/src/MapOptions.cpp:74-85, MultiIndex.cpp:243-254,
ParameterizedFunctionBase.cpp:51-56) and read them back with
``MapOptions.Deserialize``, ``FixedMultiIndexSet.Deserialize`` and
``DeserializeMap`` (bindings/python/src/Serialization.cpp:22-31).  A cereal
``BinaryOutputArchive`` stores every arithmetic value as its raw little-endian
bytes (bool as one byte, an enum as its underlying ``int``), a
``std::string`` as a uint64 length followed by its characters, and
``binary_data`` as raw bytes without a length.  The archived objects:

* MapOptions (MParT/MapOptions.h:125-131): basisType, basisLB, basisUB,
  posFuncType, quadType, quadAbsTol, quadRelTol, quadMaxSub, quadMinSub,
  quadPts, contDeriv, basisNorm, nugget.
* FixedMultiIndexSet (MParT/MultiIndices/FixedMultiIndexSet.h:113-131): dim,
  isCompressed, then the views nzStarts, nzDims, nzOrders, maxDegrees.
* A Kokkos view (MParT/Utilities/Serialization.h:19-54): its label (string),
  its length (unsigned int), then the elements.
* A map (``Serialize``): inputDim, outputDim, numCoeffs (unsigned ints), then
  the coefficient view.
"""
import struct

import numpy as np

_OPTS = [("basisType", "<i"), ("basisLB", "<d"), ("basisUB", "<d"), ("posFuncType", "<i"), ("quadType", "<i"),
         ("quadAbsTol", "<d"), ("quadRelTol", "<d"), ("quadMaxSub", "<I"), ("quadMinSub", "<I"),
         ("quadPts", "<I"), ("contDeriv", "<?"), ("basisNorm", "<?"), ("nugget", "<d")]


class _Reader:
    def __init__(self, data):
        self.b = data
        self.i = 0

    def take(self, fmt):
        v = struct.unpack_from(fmt, self.b, self.i)[0]
        self.i += struct.calcsize(fmt)
        return v

    def raw(self, n):
        if self.i + n > len(self.b):
            raise ValueError("archive ends early")
        v = self.b[self.i:self.i + n]
        self.i += n
        return v

    def string(self):
        return self.raw(self.take("<Q")).decode()

    def view(self, dtype):
        label = self.string()
        n = self.take("<I")
        arr = np.frombuffer(self.raw(n * np.dtype(dtype).itemsize), dtype=dtype).copy()
        return label, arr


def _string(s):
    b = s.encode()
    return struct.pack("<Q", len(b)) + b


def _view(label, arr, dtype):
    a = np.ascontiguousarray(arr, dtype=dtype)
    return _string(label) + struct.pack("<I", a.size) + a.astype(np.dtype(dtype).newbyteorder("<")).tobytes()


def options_to_bytes(opts):
    out = b""
    for name, fmt in _OPTS:
        v = getattr(opts, name)
        out += struct.pack(fmt, bool(v) if fmt == "<?" else (int(v) if fmt in ("<i", "<I") else float(v)))
    return out


def options_from_bytes(data, opts):
    r = _Reader(data)
    for name, fmt in _OPTS:
        setattr(opts, name, r.take(fmt))
    return opts


def multiset_to_bytes(mset):
    out = struct.pack("<I?", int(mset.dim), True)
    out += _view("nzStarts", mset.nzStarts, np.uint32)
    out += _view("nzDims", mset.nzDims, np.uint32)
    out += _view("nzOrders", mset.nzOrders, np.uint32)
    out += _view("maxDegrees", mset.MaxDegrees(), np.uint32)
    return out


def multiset_from_bytes(data):
    """(dim, nzStarts, nzDims, nzOrders) of an archived compressed set."""
    r = _Reader(data)
    dim = r.take("<I")
    compressed = r.take("<?")
    if not compressed:
        raise NotImplementedError("only compressed FixedMultiIndexSets (fix(True)) are read")
    _, starts = r.view(np.uint32)
    _, dims = r.view(np.uint32)
    _, orders = r.view(np.uint32)
    r.view(np.uint32)  # maxDegrees, recomputed from the set
    return dim, starts, dims, orders


def map_to_bytes(inputDim, outputDim, coeffs):
    c = np.asarray(coeffs, dtype=np.float64)
    return struct.pack("<III", inputDim, outputDim, c.size) + _view("Map coeffs", c, np.float64)


def map_from_bytes(data):
    """(inputDim, outputDim, coefficients) as DeserializeMap returns them."""
    r = _Reader(data)
    inputDim, outputDim, numCoeffs = r.take("<I"), r.take("<I"), r.take("<I")
    _, coeffs = r.view(np.float64)
    if coeffs.size != numCoeffs:
        raise ValueError(f"archive holds {coeffs.size} coefficients, header says {numCoeffs}")
    return inputDim, outputDim, coeffs


# ---------------------------------------------------------------------------
# A whole map as one byte string (ToBytes / FromBytes / pickle).
#
# The reference's ToBytes is cereal's polymorphic archive of the map object
# (bindings/python/src/ParameterizedFunctionBase.cpp:57-83): registered type
# names, shared-pointer ids and each class's own fields.  Those type names
# and class layouts are the reference's C++ types, which this library does
# not have, so its byte string is not reproduced.  This one is assembled from
# the archives above, which are the reference's formats:
#   b"MPARTAMD", uint32 version (1), uint8 triangular,
#   the MapOptions archive, uint32 component count,
#   per component a uint64 length and its FixedMultiIndexSet archive,
#   the map archive (inputDim, outputDim, numCoeffs, coefficient view).
_MAGIC = b"MPARTAMD"
_VERSION = 1


def map_state_to_bytes(opts, triangular, msets, inputDim, outputDim, coeffs):
    out = _MAGIC + struct.pack("<IB", _VERSION, 1 if triangular else 0)
    out += options_to_bytes(opts)
    out += struct.pack("<I", len(msets))
    for s in msets:
        b = multiset_to_bytes(s)
        out += struct.pack("<Q", len(b)) + b
    return out + map_to_bytes(inputDim, outputDim, coeffs)


def map_state_from_bytes(data, opts):
    """(triangular, [(dim, nzStarts, nzDims, nzOrders)], inputDim, outputDim,
    coefficients); fills opts."""
    data = bytes(data)
    if data[:len(_MAGIC)] != _MAGIC:
        raise ValueError("not a map byte string of this library (the reference's polymorphic cereal form is not read)")
    r = _Reader(data)
    r.raw(len(_MAGIC))
    version, tri = r.take("<I"), r.take("<B")
    if version != _VERSION:
        raise ValueError(f"map byte string version {version} is not supported")
    n_opts = struct.calcsize("<" + "".join(f[1:] for _, f in _OPTS))
    options_from_bytes(r.raw(n_opts), opts)
    sets = []
    for _ in range(r.take("<I")):
        sets.append(multiset_from_bytes(r.raw(r.take("<Q"))))
    inputDim, outputDim, coeffs = map_from_bytes(data[r.i:])
    return bool(tri), sets, inputDim, outputDim, coeffs
