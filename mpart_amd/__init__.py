"""MI355X monotone-component maps (MParT's monotone-component path on HIP).

The Python surface mirrors the reference's ``pympart`` module
(bindings/python/src/Wrapper.cpp:7-49): ``MapOptions``, ``FixedMultiIndexSet``,
``CreateComponent``, ``CreateTriangular`` and maps with ``SetCoeffs``,
``WrapCoeffs``, ``CoeffMap``, ``Evaluate``, ``LogDeterminant``, ``Gradient``,
``CoeffGrad``, ``LogDeterminantCoeffGrad``, ``LogDeterminantInputGrad`` and
``Inverse``, plus the fused ``KLValueAndGrad``.  Every call goes through the C
ABI of ``libmpart_amd.so`` (include/mpart_amd.h); torch provides device memory
and the stream.  The library must be built (``python -m mpart_amd.build``);
importing fails loudly otherwise.

Point arrays are inputDim x N (a column per point), as in the reference.
numpy inputs are copied to the GPU and results copied back; torch CUDA tensors
stay on the device.
"""
import ctypes
import math
import warnings
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmpart_amd.so")

if not os.path.exists(_LIB_PATH):
    raise ImportError(f"{_LIB_PATH} is missing: build it with `python mpart_amd/build.py`")

# torch ships its own HIP runtime; load it first so that this library binds to
# the same runtime instance that owns the tensors and streams passed to it.
import torch as _torch_rt  # noqa: E402,F401

_lib = ctypes.CDLL(_LIB_PATH)

MPART_OK, MPART_INVALID_ARGUMENT, MPART_RUNTIME_ERROR, MPART_HIP_ERROR, MPART_NOT_IMPLEMENTED = 0, -1, -2, -3, -4


class _COptions(ctypes.Structure):
    _fields_ = [
        ("basisType", ctypes.c_int32),
        ("posFuncType", ctypes.c_int32),
        ("quadType", ctypes.c_int32),
        ("contDeriv", ctypes.c_int32),
        ("basisNorm", ctypes.c_int32),
        ("quadMaxSub", ctypes.c_uint32),
        ("quadMinSub", ctypes.c_uint32),
        ("quadPts", ctypes.c_uint32),
        ("quadAbsTol", ctypes.c_double),
        ("quadRelTol", ctypes.c_double),
        ("nugget", ctypes.c_double),
        ("basisLB", ctypes.c_double),
        ("basisUB", ctypes.c_double),
    ]


class _CTrainOptions(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("stopval", ctypes.c_double),
        ("ftol_rel", ctypes.c_double),
        ("ftol_abs", ctypes.c_double),
        ("xtol_rel", ctypes.c_double),
        ("xtol_abs", ctypes.c_double),
        ("maxeval", ctypes.c_int32),
        ("maxtime", ctypes.c_double),
        ("verbose", ctypes.c_int32),
        ("history", ctypes.c_int32),
    ]


class _CView(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("rows", ctypes.c_int64), ("cols", ctypes.c_int64),
                ("row_stride", ctypes.c_int64), ("col_stride", ctypes.c_int64)]


_vp, _u32, _i64, _dbl, _int = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int64, ctypes.c_double, ctypes.c_int
_U32P = ctypes.POINTER(ctypes.c_uint32)
_OPTP = ctypes.POINTER(_COptions)
_MAPP = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "mpart_init": [],
    "mpart_total_order_set": [_u32, _u32, _U32P, _U32P, _U32P, _U32P, _U32P],
    "mpart_create_component": [_OPTP, _u32, _u32, _U32P, _U32P, _U32P, _MAPP],
    "mpart_create_total_order_component": [_OPTP, _u32, _u32, _MAPP],
    "mpart_create_triangular": [_u32, _u32, _u32, _OPTP, _MAPP],
    "mpart_create_triangular_from": [_MAPP, _u32, _MAPP],
    "mpart_destroy": [_vp],
    "mpart_map_dims": [_vp, _U32P, _U32P, _U32P],
    "mpart_map_num_components": [_vp, _U32P],
    "mpart_map_options": [_vp, _vp, ctypes.POINTER(ctypes.c_int32)],
    "mpart_component_multiset": [_vp, _u32, _U32P, _U32P, _U32P, _U32P, _U32P, _U32P],
    "mpart_set_coeffs": [_vp, _vp, _i64, _i64, _vp],
    "mpart_wrap_coeffs": [_vp, _vp, _i64, _i64],
    "mpart_get_coeffs": [_vp, _vp, _vp],
    "mpart_evaluate": [_vp, _CView, _CView, _vp],
    "mpart_evaluate_counted": [_vp, _CView, _CView, _vp, _vp],
    "mpart_log_determinant": [_vp, _CView, _vp, _i64, _vp],
    "mpart_gradient": [_vp, _CView, _CView, _CView, _vp],
    "mpart_coeff_grad": [_vp, _CView, _CView, _CView, _vp],
    "mpart_log_determinant_coeff_grad": [_vp, _CView, _CView, _vp],
    "mpart_log_determinant_input_grad": [_vp, _CView, _CView, _vp],
    "mpart_inverse": [_vp, _CView, _CView, _CView, _dbl, _dbl, _vp],
    "mpart_kl_value_and_grad": [_vp, _CView, _vp, _vp, _vp],
    "mpart_kl_value_and_grad_noalloc": [_vp, _CView, _vp, _vp, _vp],
    "mpart_codegen_kernel_info": [_vp, ctypes.c_char_p, _vp, _vp, _vp],
    "mpart_is_host_memory": [_vp],
    "mpart_check_integrand": [_vp, _vp],
    "mpart_kl_value_and_grad_plain": [_vp, _CView, _vp, _vp, _vp],
    "mpart_coeff_vjp": [_vp, _CView, _CView, _vp, _i64, _vp, _vp],
    "mpart_kl_input_grad": [_vp, _CView, _dbl, _CView, _vp],
    "mpart_input_vjp": [_vp, _CView, _CView, _vp, _i64, _CView, _vp],
    "mpart_kl_value_and_grads": [_vp, _CView, _vp, _vp, _CView, _vp],
    "mpart_vjp": [_vp, _CView, _CView, _vp, _i64, _vp, _CView, _vp],
    "mpart_set_plain_launches": [_vp, _int],
    "mpart_set_codegen": [_vp, _int],
    "mpart_create_batch": [_MAPP, _u32, _MAPP],
    "mpart_destroy_batch": [_vp],
    "mpart_batch_size": [_vp, _U32P],
    "mpart_batch_evaluate": [_vp, _vp, _vp, _vp, _vp, _vp],
    "mpart_batch_coeff_grad": [_vp, _vp, _vp, _vp, _vp],
    "mpart_batch_inverse": [_vp, _vp, _vp, _vp, _dbl, _dbl, _vp],
    "mpart_batch_set_point_group": [_vp, ctypes.c_int32],
    "mpart_batch_set_inverse_schedule": [_vp, ctypes.c_int32],
    "mpart_set_inverse_handover": [_vp, ctypes.c_int32],
    "mpart_batch_set_inverse_handover": [_vp, ctypes.c_int32],
    "mpart_inverse_handover_stats": [_vp, _vp, _vp],
    "mpart_batch_inverse_handover_stats": [_vp, _vp, _vp],
    "mpart_batch_inverse_schedule": [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
    "mpart_batch_eval_schedule": [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
    "mpart_codegen_prepare": [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
    "mpart_fused_schedule": [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
    "mpart_set_fused_schedule": [_vp, ctypes.c_int32],
    "mpart_set_point_group": [_vp, ctypes.c_int32],
    "mpart_eval_schedule": [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
    "mpart_train_map": [_vp, _CView, ctypes.POINTER(_CTrainOptions), ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_int32), _vp],
}
for _name, _args in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = ctypes.c_int
_lib.mpart_last_error.restype = ctypes.c_char_p
_lib.mpart_last_error.argtypes = []
_lib.mpart_default_options.argtypes = [_OPTP]
_lib.mpart_default_options.restype = None
_lib.mpart_codegen_error.argtypes = [_vp]
_lib.mpart_codegen_error.restype = ctypes.c_char_p
_lib.mpart_default_train_options.argtypes = [ctypes.POINTER(_CTrainOptions)]
_lib.mpart_default_train_options.restype = None

LIB_PATH = _LIB_PATH
C_EXPORTS = tuple(_SIGS) + ("mpart_last_error", "mpart_default_options", "mpart_codegen_error",
                            "mpart_default_train_options")


def _check(rc):
    if rc == MPART_OK:
        return
    msg = _lib.mpart_last_error().decode()
    if rc == MPART_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == MPART_NOT_IMPLEMENTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def Initialize():
    """Idempotent (reference: mpart::Initialize, MParT/Initialization.h:78-116)."""
    _check(_lib.mpart_init())


# ---------------------------------------------------------------------------
class BasisTypes(enum.IntEnum):
    ProbabilistHermite = 0
    PhysicistHermite = 1
    HermiteFunctions = 2


class PosFuncTypes(enum.IntEnum):
    Exp = 0
    SoftPlus = 1


class QuadTypes(enum.IntEnum):
    ClenshawCurtis = 0
    AdaptiveSimpson = 1
    AdaptiveClenshawCurtis = 2


class MapOptions:
    """Options of MParT/MapOptions.h:58-123 with the same defaults."""

    _FIELDS = ("basisType", "basisLB", "basisUB", "posFuncType", "quadType", "quadAbsTol", "quadRelTol",
               "quadMaxSub", "quadMinSub", "quadPts", "contDeriv", "basisNorm", "nugget")

    def __init__(self, **kw):
        c = _COptions()
        _lib.mpart_default_options(ctypes.byref(c))
        self.basisType = BasisTypes(c.basisType)
        self.basisLB = c.basisLB
        self.basisUB = c.basisUB
        self.posFuncType = PosFuncTypes(c.posFuncType)
        self.quadType = QuadTypes(c.quadType)
        self.quadAbsTol = c.quadAbsTol
        self.quadRelTol = c.quadRelTol
        self.quadMaxSub = c.quadMaxSub
        self.quadMinSub = c.quadMinSub
        self.quadPts = c.quadPts
        self.contDeriv = bool(c.contDeriv)
        self.basisNorm = bool(c.basisNorm)
        self.nugget = c.nugget
        for k, v in kw.items():
            if k not in self._FIELDS:
                raise ValueError(f"unknown MapOptions field {k}")
            setattr(self, k, v)

    def _c(self):
        c = _COptions()
        c.basisType = int(self.basisType)
        c.posFuncType = int(self.posFuncType)
        c.quadType = int(self.quadType)
        c.contDeriv = int(bool(self.contDeriv))
        c.basisNorm = int(bool(self.basisNorm))
        c.quadMaxSub = int(self.quadMaxSub)
        c.quadMinSub = int(self.quadMinSub)
        c.quadPts = int(self.quadPts)
        c.quadAbsTol = float(self.quadAbsTol)
        c.quadRelTol = float(self.quadRelTol)
        c.nugget = float(self.nugget)
        c.basisLB = float(self.basisLB)
        c.basisUB = float(self.basisUB)
        return c

    def __repr__(self):
        return "MapOptions(" + ", ".join(f"{k}={getattr(self, k)!r}" for k in self._FIELDS) + ")"

    def Serialize(self, filename):
        """Writes the options as the reference's cereal binary archive
        (bindings/python/src/MapOptions.cpp:74-78)."""
        from . import serialization
        with open(filename, "wb") as f:
            f.write(serialization.options_to_bytes(self))

    def Deserialize(self, filename):
        from . import serialization
        with open(filename, "rb") as f:
            serialization.options_from_bytes(f.read(), self)
        self.basisType = BasisTypes(self.basisType)
        self.posFuncType = PosFuncTypes(self.posFuncType)
        self.quadType = QuadTypes(self.quadType)
        return self


class FixedMultiIndexSet:
    """Compressed multi-index set (MParT/MultiIndices/FixedMultiIndexSet.h).

    ``FixedMultiIndexSet(dim, maxOrder)`` enumerates the total-order set in the
    reference order; ``FixedMultiIndexSet.from_csr(dim, nzStarts, nzDims, nzOrders)``
    wraps a compressed set."""

    def __init__(self, dim=None, maxOrder=None):
        self.dim = dim
        if dim is None:
            return
        nt, nz = _u32(), _u32()
        _check(_lib.mpart_total_order_set(dim, maxOrder, ctypes.byref(nt), ctypes.byref(nz), None, None, None))
        self.nzStarts = np.zeros(nt.value + 1, np.uint32)
        self.nzDims = np.zeros(nz.value, np.uint32)
        self.nzOrders = np.zeros(nz.value, np.uint32)
        _check(_lib.mpart_total_order_set(dim, maxOrder, None, None, self.nzStarts.ctypes.data_as(_U32P),
                                          self.nzDims.ctypes.data_as(_U32P), self.nzOrders.ctypes.data_as(_U32P)))

    @classmethod
    def from_csr(cls, dim, nzStarts, nzDims, nzOrders):
        s = cls()
        s.dim = int(dim)
        s.nzStarts = np.ascontiguousarray(nzStarts, np.uint32)
        s.nzDims = np.ascontiguousarray(nzDims, np.uint32)
        s.nzOrders = np.ascontiguousarray(nzOrders, np.uint32)
        return s

    def Size(self):
        return len(self.nzStarts) - 1

    def Length(self):
        return self.dim

    def MaxDegrees(self):
        out = np.zeros(self.dim, np.uint32)
        for d, o in zip(self.nzDims, self.nzOrders):
            out[d] = max(out[d], o)
        return out

    def IndexToMulti(self, i):
        m = np.zeros(self.dim, np.uint32)
        for k in range(self.nzStarts[i], self.nzStarts[i + 1]):
            m[self.nzDims[k]] = self.nzOrders[k]
        return m

    def Serialize(self, filename):
        """Writes the set as the reference's cereal binary archive
        (bindings/python/src/MultiIndex.cpp:243-248)."""
        from . import serialization
        with open(filename, "wb") as f:
            f.write(serialization.multiset_to_bytes(self))
        return self

    def Deserialize(self, filename):
        from . import serialization
        with open(filename, "rb") as f:
            self.dim, self.nzStarts, self.nzDims, self.nzOrders = serialization.multiset_from_bytes(f.read())
        self.dim = int(self.dim)
        return self


def _read_multiset(h, i):
    dim, nt, nnz = _u32(), _u32(), _u32()
    _check(_lib.mpart_component_multiset(h, i, ctypes.byref(dim), ctypes.byref(nt), ctypes.byref(nnz), None, None,
                                         None))
    starts = np.zeros(nt.value + 1, np.uint32)
    dims = np.zeros(nnz.value, np.uint32)
    orders = np.zeros(nnz.value, np.uint32)
    _check(_lib.mpart_component_multiset(h, i, None, None, None, starts.ctypes.data_as(_U32P),
                                         dims.ctypes.data_as(_U32P), orders.ctypes.data_as(_U32P)))
    return dim.value, starts, dims, orders


# ---------------------------------------------------------------------------
_HANDOVER_KEYS = ("solves", "t_evaluations", "intervals", "fallback_size", "fallback_nonfinite", "max_intervals",
                  "max_level", "max_solve_t_evaluations", "max_solve_ticks")


def _handover_stats(fn, h):
    out = (ctypes.c_uint64 * len(_HANDOVER_KEYS))()
    _check(fn(h, out, _stream()))
    return dict(zip(_HANDOVER_KEYS, (int(v) for v in out)))


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _as_dev(a):
    """(device tensor, was_numpy).  numpy arrays are copied to cuda:0."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        if a.dtype != torch.float64:
            raise ValueError(f"Currently only tensors with float64 datatype are supported, got {a.dtype}")
        if not a.is_cuda:
            return a.cuda(), True
        return a, False
    arr = np.asarray(a, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda(), True


def _view(t, rows=None):
    """Strided view of a 1-D or 2-D float64 cuda tensor."""
    if t is None:
        return _CView(None, 0, 0, 0, 0)
    if t.dim() == 1:
        return _CView(t.data_ptr(), 1, t.shape[0], 0, t.stride(0))
    return _CView(t.data_ptr(), t.shape[0], t.shape[1], t.stride(0), t.stride(1))


class ConditionalMap:
    """A monotone component or a triangular map held by libmpart_amd.so
    (the role of mpart::ConditionalMapBase<DeviceSpace>)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        i, o, c = _u32(), _u32(), _u32()
        _check(_lib.mpart_map_dims(self._h, ctypes.byref(i), ctypes.byref(o), ctypes.byref(c)))
        self.inputDim, self.outputDim, self.numCoeffs = i.value, o.value, c.value
        self._wrapped = None

    def __del__(self):
        try:
            _lib.mpart_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def NumComponents(self):
        n = _u32()
        _check(_lib.mpart_map_num_components(self._h, ctypes.byref(n)))
        return n.value

    def GetMultiIndexSet(self, i=0):
        dim, s, d, o = _read_multiset(self._h, i)
        return FixedMultiIndexSet.from_csr(dim, s, d, o)

    def Options(self):
        """(MapOptions the map was built with, whether it is a TriangularMap)."""
        c = _COptions()
        tri = ctypes.c_int32()
        _check(_lib.mpart_map_options(self._h, ctypes.byref(c), ctypes.byref(tri)))
        o = MapOptions()
        o.basisType = BasisTypes(c.basisType)
        o.posFuncType = PosFuncTypes(c.posFuncType)
        o.quadType = QuadTypes(c.quadType)
        o.contDeriv = bool(c.contDeriv)
        o.basisNorm = bool(c.basisNorm)
        o.quadMaxSub, o.quadMinSub, o.quadPts = c.quadMaxSub, c.quadMinSub, c.quadPts
        o.quadAbsTol, o.quadRelTol, o.nugget = c.quadAbsTol, c.quadRelTol, c.nugget
        o.basisLB, o.basisUB = c.basisLB, c.basisUB
        return o, bool(tri.value)

    # byte string and pickle (ParameterizedFunctionBase ToBytes / FromBytes and
    # py::pickle, bindings/python/src/ParameterizedFunctionBase.cpp:57-83) ------
    def ToBytes(self):
        """The map (options, multi-index sets, coefficients) as one byte string
        (mpart_amd.serialization.map_state_to_bytes; not the reference's
        polymorphic cereal bytes)."""
        from . import serialization

        opts, tri = self.Options()
        msets = [self.GetMultiIndexSet(i) for i in range(self.NumComponents())]
        coeffs = self.CoeffMap()
        return serialization.map_state_to_bytes(opts, tri, msets, self.inputDim, self.outputDim, coeffs)

    @staticmethod
    def FromBytes(data):
        """A new map from ToBytes' byte string."""
        from . import serialization

        opts = MapOptions()
        tri, sets, inputDim, outputDim, coeffs = serialization.map_state_from_bytes(data, opts)
        opts.basisType = BasisTypes(opts.basisType)
        opts.posFuncType = PosFuncTypes(opts.posFuncType)
        opts.quadType = QuadTypes(opts.quadType)
        comps = [CreateComponent(FixedMultiIndexSet.from_csr(*st), opts) for st in sets]
        m = TriangularMap(comps) if tri else comps[0]
        if (m.inputDim, m.outputDim, m.numCoeffs) != (inputDim, outputDim, coeffs.size):
            raise ValueError("byte string: the multi-index sets do not match the archived dimensions")
        m.SetCoeffs(coeffs)
        return m

    def __reduce__(self):
        return (ConditionalMap.FromBytes, (self.ToBytes(),))

    # coefficients ---------------------------------------------------------
    def SetCoeffs(self, coeffs):
        """Deep copy (src/ParameterizedFunctionBase.cpp:184-214)."""
        t, _ = _as_dev(coeffs)
        if t.dim() != 1:
            raise ValueError("coefficients must be a vector")
        _check(_lib.mpart_set_coeffs(self._h, t.data_ptr(), t.shape[0], t.stride(0), _stream()))
        _torch().cuda.current_stream().synchronize()
        self._wrapped = None

    def WrapCoeffs(self, coeffs):
        """Alias a float64 cuda vector without copying (219-228)."""
        torch = _torch()
        if not (isinstance(coeffs, torch.Tensor) and coeffs.is_cuda and coeffs.dtype == torch.float64):
            raise ValueError("WrapCoeffs needs a float64 cuda tensor")
        _check(_lib.mpart_wrap_coeffs(self._h, coeffs.data_ptr(), coeffs.shape[0], coeffs.stride(0)))
        self._wrapped = coeffs

    def torch(self, store_coeffs=True, return_logdet=False):
        """torch.nn.Module view of the map (bindings/python/src/ConditionalMapBase.cpp:44-50)."""
        from .torch import TorchConditionalMapBase

        return TorchConditionalMapBase(self, store_coeffs, return_logdet)

    def Serialize(self, filename):
        """inputDim, outputDim, numCoeffs and the coefficients as the
        reference's cereal binary archive (bindings/python/src/ParameterizedFunctionBase.cpp:51-56)."""
        from . import serialization
        with open(filename, "wb") as f:
            f.write(serialization.map_to_bytes(self.inputDim, self.outputDim, self.CoeffMap()))

    def CoeffMap(self):
        torch = _torch()
        out = torch.empty(self.numCoeffs, dtype=torch.float64, device="cuda")
        _check(_lib.mpart_get_coeffs(self._h, out.data_ptr(), _stream()))
        return out.cpu().numpy()

    def SetPlainLaunches(self, plain=True):
        _check(_lib.mpart_set_plain_launches(self._h, int(bool(plain))))

    def SetCodegen(self, enable=True):
        """Specialized (hipRTC) kernels when enabled (default), generic kernels otherwise."""
        _check(_lib.mpart_set_codegen(self._h, int(bool(enable))))

    def PrepareCodegen(self):
        """Compiles the specialized kernels now; returns (specialized, compile_seconds, error)."""
        spec = ctypes.c_int32(0)
        secs = ctypes.c_double(0.0)
        _check(_lib.mpart_codegen_prepare(self._h, ctypes.byref(spec), ctypes.byref(secs)))
        return bool(spec.value), float(secs.value), _lib.mpart_codegen_error(self._h).decode()

    def SetInverseHandover(self, solveSteps=1024):
        """At the end of an Inverse pass, solves longer than solveSteps quadrature interval steps are
        finished by a wavefront each (``mpart_set_inverse_handover``); 0 turns it off."""
        _check(_lib.mpart_set_inverse_handover(self._h, int(solveSteps)))

    def InverseHandoverStats(self):
        """Totals of the last Inverse call with the hand-over on (``mpart_inverse_handover_stats``)."""
        return _handover_stats(_lib.mpart_inverse_handover_stats, self._h)

    def SetPointGroup(self, lanes=0):
        """Lanes per point of the generic Evaluate / LogDeterminant / fused pass (1, 4, 8, 64; 0 = autotune;
        2 = the components split over the wavefronts of a block)."""
        _check(_lib.mpart_set_point_group(self._h, int(lanes)))

    def EvalSchedule(self):
        """((lanes for Evaluate, LogDeterminant, fused pass), trial ms of the generic kernels for each: lanes
        1, 4, 8, 64, then the split schedule (code 2)) (0: not tuned)."""
        lanes = (ctypes.c_int32 * 3)()
        ms = (ctypes.c_double * 15)()
        _check(_lib.mpart_eval_schedule(self._h, lanes, ms))
        return (lanes[0], lanes[1], lanes[2]), [list(ms[0:5]), list(ms[5:10]), list(ms[10:15])]

    def KernelInfo(self, kernel="fused"):
        """(registers, scratch bytes, LDS bytes) of a compiled specialized kernel."""
        v, sc, lds = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.mpart_codegen_kernel_info(self._h, kernel.encode(), ctypes.byref(v), ctypes.byref(sc),
                                              ctypes.byref(lds)))
        return int(v.value), int(sc.value), int(lds.value)

    def SetFusedSchedule(self, choice=-1):
        """Forces the specialized fused kernel's schedule (0-3, mpart_set_fused_schedule) or -1 to tune again."""
        _check(_lib.mpart_set_fused_schedule(self._h, int(choice)))

    def FusedSchedule(self):
        """(schedule index or -1 before tuning, trial times in ms) of the fused kernel."""
        c = ctypes.c_int32(-1)
        t = (ctypes.c_double * 4)()
        _check(_lib.mpart_fused_schedule(self._h, ctypes.byref(c), t))
        return int(c.value), list(t)

    # Impl methods (device tensors, preallocated outputs) --------------------
    def EvaluateImpl(self, pts, out):
        _check(_lib.mpart_evaluate(self._h, _view(pts), _view(out), _stream()))

    def LogDeterminantImpl(self, pts, out):
        _check(_lib.mpart_log_determinant(self._h, _view(pts), out.data_ptr(), out.stride(0), _stream()))

    def GradientImpl(self, pts, sens, out):
        _check(_lib.mpart_gradient(self._h, _view(pts), _view(sens), _view(out), _stream()))

    def CoeffGradImpl(self, pts, sens, out):
        _check(_lib.mpart_coeff_grad(self._h, _view(pts), _view(sens), _view(out), _stream()))

    def LogDeterminantCoeffGradImpl(self, pts, out):
        _check(_lib.mpart_log_determinant_coeff_grad(self._h, _view(pts), _view(out), _stream()))

    def LogDeterminantInputGradImpl(self, pts, out):
        _check(_lib.mpart_log_determinant_input_grad(self._h, _view(pts), _view(out), _stream()))

    def InverseImpl(self, x1, r, out, xtol=1e-6, ytol=1e-6):
        _check(_lib.mpart_inverse(self._h, _view(x1), _view(r), _view(out), xtol, ytol, _stream()))

    def CheckIntegrand(self):
        """Raises the reference's RuntimeError (MonotoneIntegrand.h:183-190) when a
        call on this map met an infinite g(df) at a quadrature node since the
        last check, and clears the record; waits for the current stream.  The
        methods below check by themselves when they return numpy arrays
        (synchronous calls); with device tensors the check is left to the
        caller, as the calls are asynchronous."""
        _check(_lib.mpart_check_integrand(self._h, _stream()))

    # allocating methods -----------------------------------------------------
    def _run(self, fn, shape, *ins):
        torch = _torch()
        devs = [_as_dev(a) for a in ins]
        host = any(h for _, h in devs)
        out = torch.zeros(shape, dtype=torch.float64, device="cuda")
        fn(*[t for t, _ in devs], out)
        if host:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return out.cpu().numpy()
        return out

    def Evaluate(self, pts):
        return self._run(self.EvaluateImpl, (self.outputDim, _ncols(pts)), pts)

    def EvaluateCounted(self, pts):
        """Evaluate plus the number of integrand evaluations per (component, point)."""
        torch = _torch()
        p, host = _as_dev(pts)
        out = torch.zeros((self.outputDim, p.shape[1]), dtype=torch.float64, device="cuda")
        ne = torch.zeros((self.NumComponents(), p.shape[1]), dtype=torch.int32, device="cuda")
        _check(_lib.mpart_evaluate_counted(self._h, _view(p), _view(out), ne.data_ptr(), _stream()))
        if host:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return out.cpu().numpy(), ne.cpu().numpy()
        return out, ne

    def LogDeterminant(self, pts):
        return self._run(self.LogDeterminantImpl, (_ncols(pts),), pts)

    def Gradient(self, pts, sens):
        return self._run(self.GradientImpl, (self.inputDim, _ncols(pts)), pts, sens)

    def CoeffGrad(self, pts, sens):
        return self._run(self.CoeffGradImpl, (self.numCoeffs, _ncols(pts)), pts, sens)

    def LogDeterminantCoeffGrad(self, pts):
        return self._run(self.LogDeterminantCoeffGradImpl, (self.numCoeffs, _ncols(pts)), pts)

    def LogDeterminantInputGrad(self, pts):
        return self._run(self.LogDeterminantInputGradImpl, (self.inputDim, _ncols(pts)), pts)

    def Inverse(self, x1, r, xtol=1e-6, ytol=1e-6):
        torch = _torch()
        rr, host_r = _as_dev(r)
        if rr.dim() == 1:
            rr = rr[None, :]
        if x1 is None or (hasattr(x1, "shape") and x1.shape[0] == 0):
            xx, host_x = torch.zeros((0, rr.shape[1]), dtype=torch.float64, device="cuda"), False
        else:
            xx, host_x = _as_dev(x1)
        out = torch.zeros((self.outputDim, rr.shape[1]), dtype=torch.float64, device="cuda")
        self.InverseImpl(xx, rr, out, xtol, ytol)
        if host_r or host_x:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return out.cpu().numpy()
        return out

    def KLValueAndGradImpl(self, pts, loss, grad, plain=False):
        """KLValueAndGrad into preallocated device outputs (loss: 1 double, grad: numCoeffs)."""
        fn = _lib.mpart_kl_value_and_grad_plain if plain else _lib.mpart_kl_value_and_grad
        _check(fn(self._h, _view(pts), loss.data_ptr(), grad.data_ptr(), _stream()))

    def KLValueAndGrad(self, pts, plain=False):
        """Fused forward + adjoint of mean(0.5 |T|^2 - logdet) and its coefficient gradient."""
        torch = _torch()
        p, host = _as_dev(pts)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        grad = torch.zeros(self.numCoeffs, dtype=torch.float64, device="cuda")
        fn = _lib.mpart_kl_value_and_grad_plain if plain else _lib.mpart_kl_value_and_grad
        _check(fn(self._h, _view(p), loss.data_ptr(), grad.data_ptr(), _stream()))
        if host:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return float(loss.item()), grad.cpu().numpy()
        return loss, grad

    def KLValueAndGradNoAllocImpl(self, pts, loss, grad):
        """KLValueAndGradImpl with every work buffer in the library's fixed
        per-GPU scratch: allocates nothing (no reference counterpart)."""
        _check(_lib.mpart_kl_value_and_grad_noalloc(self._h, _view(pts), loss.data_ptr(), grad.data_ptr(),
                                                    _stream()))

    def KLValueAndGradsImpl(self, pts, loss, grad, grad_inputs):
        """Loss, coefficient gradient and (grad_inputs, inputDim x N, may be None)
        the per-point input gradient of the mean KL objective, from one pass."""
        _check(_lib.mpart_kl_value_and_grads(self._h, _view(pts), loss.data_ptr(), grad.data_ptr(),
                                             _view(grad_inputs), _stream()))

    def VJPImpl(self, pts, sens, logdet_sens, grad, grad_inputs):
        """CoeffVJP into grad and InputVJP into grad_inputs (may be None), from one pass."""
        ls_ptr = logdet_sens.data_ptr() if logdet_sens is not None else None
        ls_stride = logdet_sens.stride(0) if logdet_sens is not None else 0
        _check(_lib.mpart_vjp(self._h, _view(pts), _view(sens), ls_ptr, ls_stride, grad.data_ptr(),
                              _view(grad_inputs), _stream()))

    def KLInputGradImpl(self, pts, out, scale=1.0):
        """out (inputDim x N) = scale * grad_x [0.5 |T(x)|^2 - log det grad T(x)], per point, fused."""
        _check(_lib.mpart_kl_input_grad(self._h, _view(pts), float(scale), _view(out), _stream()))

    def KLInputGrad(self, pts, scale=1.0):
        """Gradient(pts, T(pts)) - LogDeterminantInputGrad(pts), times scale, in one launch."""
        torch = _torch()
        p, host = _as_dev(pts)
        out = torch.empty((self.inputDim, p.shape[1]), dtype=torch.float64, device="cuda")
        self.KLInputGradImpl(p, out, scale)
        if host:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return out.cpu().numpy()
        return out

    def InputVJPImpl(self, pts, sens, logdet_sens, out):
        """out (inputDim x N) = Gradient(sens) + logdet_sens[n] LogDeterminantInputGrad, fused
        (sens and logdet_sens may be None)."""
        ls_ptr = logdet_sens.data_ptr() if logdet_sens is not None else None
        ls_stride = logdet_sens.stride(0) if logdet_sens is not None else 0
        _check(_lib.mpart_input_vjp(self._h, _view(pts), _view(sens), ls_ptr, ls_stride, _view(out), _stream()))

    def InputVJP(self, pts, sens=None, logdet_sens=None):
        torch = _torch()
        p, host = _as_dev(pts)
        s = _as_dev(sens)[0] if sens is not None else None
        ls = _as_dev(logdet_sens)[0].contiguous() if logdet_sens is not None else None
        out = torch.empty((self.inputDim, p.shape[1]), dtype=torch.float64, device="cuda")
        self.InputVJPImpl(p, s, ls, out)
        if host:
            torch.cuda.current_stream().synchronize()
            self.CheckIntegrand()
            return out.cpu().numpy()
        return out

    def CoeffVJP(self, pts, sens=None, logdet_sens=None, out=None):
        """sum_n CoeffGrad(sens)(:, n) + logdet_sens[n] LogDeterminantCoeffGrad(:, n), fused."""
        torch = _torch()
        if out is None:
            out = torch.zeros(self.numCoeffs, dtype=torch.float64, device="cuda")
        ls_ptr = logdet_sens.data_ptr() if logdet_sens is not None else None
        ls_stride = logdet_sens.stride(0) if logdet_sens is not None else 0
        _check(_lib.mpart_coeff_vjp(self._h, _view(pts), _view(sens), ls_ptr, ls_stride, out.data_ptr(), _stream()))
        return out


def _ncols(a):
    return a.shape[1]


# ---------------------------------------------------------------------------
def CreateComponent(mset, opts=None):
    """MapFactory::CreateComponent (src/MapFactory.cpp:22-27)."""
    opts = opts or MapOptions()
    h = ctypes.c_void_p()
    s = np.ascontiguousarray(mset.nzStarts, np.uint32)
    d = np.ascontiguousarray(mset.nzDims, np.uint32)
    o = np.ascontiguousarray(mset.nzOrders, np.uint32)
    _check(_lib.mpart_create_component(ctypes.byref(opts._c()), mset.dim, len(s) - 1, s.ctypes.data_as(_U32P),
                                       d.ctypes.data_as(_U32P), o.ctypes.data_as(_U32P), ctypes.byref(h)))
    return ConditionalMap(h.value)


def CreateTotalOrderComponent(dim, order, opts=None):
    opts = opts or MapOptions()
    h = ctypes.c_void_p()
    _check(_lib.mpart_create_total_order_component(ctypes.byref(opts._c()), dim, order, ctypes.byref(h)))
    return ConditionalMap(h.value)


def CreateTriangular(inputDim, outputDim, totalOrder, opts=None):
    """MapFactory::CreateTriangular (src/MapFactory.cpp:105-125)."""
    opts = opts or MapOptions()
    h = ctypes.c_void_p()
    _check(_lib.mpart_create_triangular(inputDim, outputDim, totalOrder, ctypes.byref(opts._c()), ctypes.byref(h)))
    return ConditionalMap(h.value)


def TriangularMap(components, moveCoeffs=False):
    """TriangularMap(components, moveCoeffs) (src/TriangularMap.cpp:10-61): the
    components' coefficients are copied into the new map either way."""
    arr = (ctypes.c_void_p * len(components))(*[c.handle.value for c in components])
    h = ctypes.c_void_p()
    _check(_lib.mpart_create_triangular_from(arr, len(components), ctypes.byref(h)))
    return ConditionalMap(h.value)


TriangularMap.FromBytes = ConditionalMap.FromBytes


class MapBatch:
    """Many independent maps (sharing MapOptions) run one method in one
    persistent launch (``mpart_create_batch``).  Arguments are lists with one
    device tensor per map; outputs are preallocated (``*Impl``) or allocated."""

    def __init__(self, maps):
        self.maps = list(maps)
        arr = (ctypes.c_void_p * len(self.maps))(*[m.handle.value for m in self.maps])
        h = ctypes.c_void_p()
        _check(_lib.mpart_create_batch(arr, len(self.maps), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        # at interpreter exit the module globals may already be gone
        if h is not None and h.value and _lib is not None:
            _lib.mpart_destroy_batch(h)
            self._h = None

    def SetPointGroup(self, lanes=0):
        """Batched Evaluate / LogDeterminant schedule: 1 (a lane per point), 4 (maps with many
        terms take 4 lanes per point), 0 = autotune."""
        _check(_lib.mpart_batch_set_point_group(self._h, int(lanes)))

    def SetInverseSchedule(self, schedule=0):
        """Batched Inverse schedule: 1 per-lane solve scheduling over every map's
        points, component by component; 2 a lane per point; 0 autotune."""
        _check(_lib.mpart_batch_set_inverse_schedule(self._h, int(schedule)))

    def SetInverseHandover(self, solveSteps=1024):
        """ConditionalMap.SetInverseHandover for the batch's schedule-1 Inverse."""
        _check(_lib.mpart_batch_set_inverse_handover(self._h, int(solveSteps)))

    def InverseHandoverStats(self):
        """ConditionalMap.InverseHandoverStats for the batch."""
        return _handover_stats(_lib.mpart_batch_inverse_handover_stats, self._h)

    def InverseSchedule(self):
        """(schedule: 1, 2, or 0 before tuning; [trial ms of schedule 1, of schedule 2])."""
        sched = ctypes.c_int32(0)
        ms = (ctypes.c_double * 2)()
        _check(_lib.mpart_batch_inverse_schedule(self._h, ctypes.byref(sched), ms))
        return int(sched.value), list(ms)

    def EvalSchedule(self):
        """(schedule: 1, 4, or 0 before tuning; [trial ms of schedule 1, of schedule 4])."""
        lanes = ctypes.c_int32(0)
        ms = (ctypes.c_double * 2)()
        _check(_lib.mpart_batch_eval_schedule(self._h, ctypes.byref(lanes), ms))
        return int(lanes.value), list(ms)

    def __len__(self):
        return len(self.maps)

    def CheckIntegrand(self):
        """ConditionalMap.CheckIntegrand for every map of the batch."""
        for m in self.maps:
            m.CheckIntegrand()

    @staticmethod
    def _views(ts):
        return (_CView * len(ts))(*[_view(t) for t in ts])

    def EvaluateImpl(self, pts, outs=None, logdets=None):
        pv = self._views(pts)
        ov = self._views(outs) if outs is not None else None
        if logdets is not None:
            lp = (ctypes.c_void_p * len(logdets))(*[t.data_ptr() for t in logdets])
            ls = (ctypes.c_int64 * len(logdets))(*[t.stride(0) for t in logdets])
        else:
            lp = ls = None
        _check(_lib.mpart_batch_evaluate(self._h, ctypes.cast(pv, _vp), ctypes.cast(ov, _vp) if ov is not None else None,
                                         ctypes.cast(lp, _vp) if lp is not None else None,
                                         ctypes.cast(ls, _vp) if ls is not None else None, _stream()))

    def CoeffGradImpl(self, pts, sens, outs):
        _check(_lib.mpart_batch_coeff_grad(self._h, ctypes.cast(self._views(pts), _vp), ctypes.cast(self._views(sens), _vp),
                                           ctypes.cast(self._views(outs), _vp), _stream()))

    def InverseImpl(self, x1, r, outs, xtol=1e-6, ytol=1e-6):
        _check(_lib.mpart_batch_inverse(self._h, ctypes.cast(self._views(x1), _vp), ctypes.cast(self._views(r), _vp),
                                        ctypes.cast(self._views(outs), _vp), xtol, ytol, _stream()))

    def Evaluate(self, pts, logdet=False):
        torch = _torch()
        outs = [torch.empty((m.outputDim, p.shape[1]), dtype=torch.float64, device=p.device) for m, p in zip(self.maps, pts)]
        lds = [torch.empty(p.shape[1], dtype=torch.float64, device=p.device) for p in pts] if logdet else None
        self.EvaluateImpl(pts, outs, lds)
        return (outs, lds) if logdet else outs

    def LogDeterminant(self, pts):
        torch = _torch()
        lds = [torch.empty(p.shape[1], dtype=torch.float64, device=p.device) for p in pts]
        self.EvaluateImpl(pts, None, lds)
        return lds

    def CoeffGrad(self, pts, sens):
        torch = _torch()
        outs = [torch.empty((m.numCoeffs, p.shape[1]), dtype=torch.float64, device=p.device) for m, p in zip(self.maps, pts)]
        self.CoeffGradImpl(pts, sens, outs)
        return outs

    def Inverse(self, x1, r, xtol=1e-6, ytol=1e-6):
        torch = _torch()
        outs = [torch.empty((m.outputDim, rr.shape[1]), dtype=torch.float64, device=rr.device) for m, rr in zip(self.maps, r)]
        self.InverseImpl(x1, r, outs, xtol, ytol)
        return outs


__all__ = ["MapOptions", "BasisTypes", "PosFuncTypes", "QuadTypes", "FixedMultiIndexSet", "ConditionalMap",
           "CreateComponent", "CreateTotalOrderComponent", "CreateTriangular", "TriangularMap", "MapBatch",
           "Initialize"]


# ---------------------------------------------------------------------------
# Training (TrainOptions, MParT/TrainMap.h:19-66; CreateGaussianKLObjective,
# src/MapObjective.cpp:41-52; TrainMap, src/TrainMap.cpp:168-216).

_TRAIN_ALGS = ("LD_LBFGS", "LD_SLSQP")


class TrainOptions:
    """Fields and defaults of the reference's TrainOptions.  The optimizer is
    L-BFGS with a strong-Wolfe line search; opt_alg accepts "LD_LBFGS" and
    the reference's default "LD_SLSQP" and rejects other NLopt names.  NLopt's
    SLSQP is not provided: "LD_SLSQP" runs L-BFGS (another quasi-Newton method
    on the same unconstrained problem, so the optimum agrees to the stopping
    tolerances but the iterates differ), and training with it issues a
    RuntimeWarning that says so.  ``history`` (L-BFGS memory) has no reference
    counterpart."""

    def __init__(self, **kw):
        self.opt_alg = "LD_SLSQP"
        self.opt_stopval = -math.inf
        self.opt_ftol_rel = 1e-3
        self.opt_ftol_abs = 1e-3
        self.opt_xtol_rel = 1e-4
        self.opt_xtol_abs = 1e-4
        self.opt_maxeval = 1000
        self.opt_maxtime = math.inf
        self.verbose = 0
        self.history = 10
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(f"TrainOptions has no field {k}")
            setattr(self, k, v)

    def String(self):
        return "\n".join(f"{k} = {getattr(self, k)}" for k in (
            "opt_alg", "opt_stopval", "opt_ftol_rel", "opt_ftol_abs", "opt_xtol_rel", "opt_xtol_abs",
            "opt_maxeval", "opt_maxtime", "verbose"))

    __str__ = String

    def _c(self):
        if self.opt_alg not in _TRAIN_ALGS:
            raise NotImplementedError(f"TrainOptions: optimizer {self.opt_alg!r} is not available (L-BFGS only)")
        if self.opt_alg == "LD_SLSQP":
            warnings.warn("TrainOptions: opt_alg 'LD_SLSQP' (NLopt's SLSQP) is not provided; training runs L-BFGS "
                          "instead (set opt_alg='LD_LBFGS' to silence this)", RuntimeWarning, stacklevel=3)
        o = _CTrainOptions()
        _lib.mpart_default_train_options(ctypes.byref(o))
        o.stopval, o.ftol_rel, o.ftol_abs = self.opt_stopval, self.opt_ftol_rel, self.opt_ftol_abs
        o.xtol_rel, o.xtol_abs = self.opt_xtol_rel, self.opt_xtol_abs
        o.maxeval, o.maxtime, o.verbose, o.history = int(self.opt_maxeval), self.opt_maxtime, int(self.verbose), \
            int(self.history)
        return o


class KLObjective:
    """KL divergence to the standard Gaussian (KLObjective with a
    GaussianSamplerDensity, src/MapObjective.cpp:41-91): the objective of a map
    on samples X is mean_n(|T(x_n)|^2 / 2 + d log(2 pi) / 2 - log det dT(x_n)),
    d = the map's output dimension.  Samples are kept on the device."""

    def __init__(self, train, test=None, dim=0):
        self.train, _ = _as_dev(train)
        self.test = _as_dev(test)[0] if test is not None else None
        self.dim = int(dim) if dim else int(self.train.shape[0])

    def _value(self, X, map):
        loss, _ = map.KLValueAndGrad(X)
        v = float(loss.item())
        map.CheckIntegrand()
        return v + 0.5 * map.outputDim * math.log(2.0 * math.pi)

    def TrainError(self, map):
        return self._value(self.train, map)

    def TestError(self, map):
        if self.test is None:
            raise RuntimeError("No test dataset given!")
        return self._value(self.test, map)

    def ObjectivePlusCoeffGrad(self, map, X=None):
        """(objective, coefficient gradient) on X (default: the training samples)."""
        X = self.train if X is None else _as_dev(X)[0]
        loss, grad = map.KLValueAndGrad(X)
        v = float(loss.item())
        map.CheckIntegrand()
        return v + 0.5 * map.outputDim * math.log(2.0 * math.pi), grad


def CreateGaussianKLObjective(train, test=None, dim=0):
    """ObjectiveFactory::CreateGaussianKLObjective(train[, test], dim)."""
    if isinstance(test, int) and dim == 0:
        test, dim = None, test
    return KLObjective(train, test, dim)


def TrainMap(map, objective, options=None):
    """Trains `map` in place on the objective's training samples
    (src/TrainMap.cpp:168-216).  A triangular map is trained one component at
    a time and the sum of the per-component objective values is returned
    (without the Gaussian normalization constant, as the reference's
    FastGaussianReverseKLObjective gives them); a lone component is trained as
    a whole and the KL objective's value (with 0.5 d log(2 pi)) is returned."""
    if not isinstance(objective, KLObjective):
        raise NotImplementedError("TrainMap: only the Gaussian KL objective is available")
    if objective.dim != map.outputDim:
        raise ValueError(f"TrainMap: objective dimension {objective.dim} != map output dimension {map.outputDim}")
    opts = (options or TrainOptions())._c()
    err, nev = ctypes.c_double(), ctypes.c_int32()
    _check(_lib.mpart_train_map(map.handle, _view(objective.train), ctypes.byref(opts), ctypes.byref(err),
                                ctypes.byref(nev), _stream()))
    map.last_train_evaluations = nev.value
    return err.value


def DeserializeMap(filename):
    """(inputDim, outputDim, coefficients) from a file written by Serialize
    (bindings/python/src/Serialization.cpp:22-31)."""
    from . import serialization
    with open(filename, "rb") as f:
        return serialization.map_from_bytes(f.read())
