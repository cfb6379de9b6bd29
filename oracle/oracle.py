"""ctypes front end of the CPU oracle (oracle/mpart_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker.  The
product package ``mpart_amd`` never imports this module.

Arrays are numpy float64.  Point matrices are D x N (a column per point), as in
the reference's ``StridedMatrix`` convention; any strides are accepted.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmpart_oracle.so")

BASIS = {"ProbabilistHermite": 0, "PhysicistHermite": 1, "HermiteFunctions": 2}
POSFUNC = {"Exp": 0, "SoftPlus": 1}
QUAD = {"ClenshawCurtis": 0, "AdaptiveSimpson": 1, "AdaptiveClenshawCurtis": 2}


class COptions(ctypes.Structure):
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


def options(basisType="ProbabilistHermite", posFuncType="SoftPlus", quadType="AdaptiveSimpson",
            quadAbsTol=1e-6, quadRelTol=1e-6, quadMaxSub=30, quadMinSub=0, quadPts=5,
            contDeriv=True, basisNorm=True, nugget=0.0, basisLB=-math.inf, basisUB=math.inf):
    """MapOptions with the reference defaults (MParT/MapOptions.h:61-123)."""
    o = COptions()
    o.basisType = BASIS[basisType]
    o.posFuncType = POSFUNC[posFuncType]
    o.quadType = QUAD[quadType]
    o.contDeriv = int(bool(contDeriv))
    o.basisNorm = int(bool(basisNorm))
    o.quadMaxSub = quadMaxSub
    o.quadMinSub = quadMinSub
    o.quadPts = quadPts
    o.quadAbsTol = quadAbsTol
    o.quadRelTol = quadRelTol
    o.nugget = nugget
    o.basisLB = basisLB
    o.basisUB = basisUB
    return o


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_last_error.restype = ctypes.c_char_p
        _declare(_lib)
    return _lib


def _declare(L):
    c = ctypes
    vp, D, u, l, i, d = c.c_void_p, c.POINTER(c.c_double), c.c_uint, c.c_long, c.c_int, c.c_double
    U, Lp, Ip, Dp = c.POINTER(c.c_uint32), c.POINTER(c.c_long), c.POINTER(c.c_int), c.POINTER(c.c_double)
    sig = {
        "orc_set_threads": [i],
        "orc_total_order_size": [u, u, c.POINTER(u), c.POINTER(u)],
        "orc_total_order_fill": [u, u, U, U, U, U],
        "orc_basis_eval": [i, i, i, u, d, D, D, D],
        "orc_basis_eval_lin": [i, i, i, u, d, d, d, D, D, D],
        "orc_posfunc_eval": [i, d, Dp, Dp, Dp],
        "orc_cc_rule": [u, D, D],
        "orc_quad_test": [i, u, u, d, d, i, d, d, Dp, Lp],
        "orc_itp_test": [i, d, d, d, d, Ip, Dp],
        "orc_find_bracket_test": [i, d, Dp, Dp, Dp, Dp, Ip],
        "orc_create_component": [vp, u, u, U, U, U, c.POINTER(vp)],
        "orc_create_total_order_component": [vp, u, u, c.POINTER(vp)],
        "orc_create_triangular": [vp, u, u, u, c.POINTER(vp)],
        "orc_create_triangular_orders": [vp, u, c.POINTER(u), c.POINTER(u), c.POINTER(vp)],
        "orc_destroy": [vp],
        "orc_info": [vp, c.POINTER(u), c.POINTER(u), c.POINTER(u)],
        "orc_set_coeffs": [vp, D],
        "orc_evaluate": [vp, D, l, l, l, l, D, l, l, Lp],
        "orc_log_determinant": [vp, D, l, l, l, l, D],
        "orc_coeff_grad": [vp, D, l, l, l, l, D, l, l, D, l, l],
        "orc_log_determinant_coeff_grad": [vp, D, l, l, l, l, D, l, l],
        "orc_gradient": [vp, D, l, l, l, l, D, l, l, D, l, l],
        "orc_log_determinant_input_grad": [vp, D, l, l, l, l, D, l, l],
        "orc_inverse": [vp, D, l, l, l, l, D, l, l, l, l, D, l, l, d, d],
        "orc_kl": [vp, D, l, l, l, l, i, Dp, D],
    }
    for name, args in sig.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = c.c_int
    L.orc_destroy.restype = None


_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)
_L = ctypes.POINTER(ctypes.c_long)


def _dp(a):
    return a.ctypes.data_as(_D)


def _strides(a):
    return a.strides[0] // 8, a.strides[1] // 8


def _check(rc):
    if rc != 0:
        msg = lib().orc_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        raise RuntimeError(msg)


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


def total_order(dim, order):
    L = lib()
    nt, nz = ctypes.c_uint(), ctypes.c_uint()
    L.orc_total_order_size(dim, order, ctypes.byref(nt), ctypes.byref(nz))
    starts = np.zeros(nt.value + 1, np.uint32)
    dims = np.zeros(nz.value, np.uint32)
    orders = np.zeros(nz.value, np.uint32)
    maxdeg = np.zeros(dim, np.uint32)
    L.orc_total_order_fill(dim, order, starts.ctypes.data_as(_U), dims.ctypes.data_as(_U),
                           orders.ctypes.data_as(_U), maxdeg.ctypes.data_as(_U))
    return starts, dims, orders, maxdeg


def basis_eval(basis, x, maxOrder, mode=2, normalize=True):
    v = np.zeros(maxOrder + 1)
    d = np.zeros(maxOrder + 1)
    d2 = np.zeros(maxOrder + 1)
    lib().orc_basis_eval(BASIS[basis], int(normalize), mode, maxOrder, ctypes.c_double(x), _dp(v), _dp(d), _dp(d2))
    return v, d, d2


def basis_eval_lin(basis, x, maxOrder, lb, ub, mode=2, normalize=True):
    """LinearizedBasis(basis, lb, ub) values / derivatives / second derivatives at x."""
    v = np.zeros(maxOrder + 1)
    d = np.zeros(maxOrder + 1)
    d2 = np.zeros(maxOrder + 1)
    lib().orc_basis_eval_lin(BASIS[basis], int(normalize), mode, maxOrder, ctypes.c_double(x), ctypes.c_double(lb),
                             ctypes.c_double(ub), _dp(v), _dp(d), _dp(d2))
    return v, d, d2


def posfunc_eval(name, x):
    g, dg, d2g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_posfunc_eval(POSFUNC[name], ctypes.c_double(x), ctypes.byref(g), ctypes.byref(dg), ctypes.byref(d2g))
    return g.value, dg.value, d2g.value


def cc_rule(n):
    w = np.zeros(n)
    p = np.zeros(n)
    lib().orc_cc_rule(n, _dp(w), _dp(p))
    return w, p


def quad_test(quadType, level, maxSub, absTol, relTol, mode, lb=0.0, ub=1.0):
    r = ctypes.c_double()
    ne = ctypes.c_long()
    lib().orc_quad_test(QUAD[quadType], level, maxSub, ctypes.c_double(absTol), ctypes.c_double(relTol), mode,
                        ctypes.c_double(lb), ctypes.c_double(ub), ctypes.byref(r), ctypes.byref(ne))
    return r.value, ne.value


def itp_test(fn, yd, x0, xtol, ftol):
    info = ctypes.c_int()
    x = ctypes.c_double()
    lib().orc_itp_test(fn, ctypes.c_double(yd), ctypes.c_double(x0), ctypes.c_double(xtol), ctypes.c_double(ftol),
                       ctypes.byref(info), ctypes.byref(x))
    return x.value, info.value


def find_bracket_test(fn, yd, xlb, ylb, xub, yub):
    a, b, c, d = (ctypes.c_double(v) for v in (xlb, ylb, xub, yub))
    info = ctypes.c_int()
    lib().orc_find_bracket_test(fn, ctypes.c_double(yd), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                ctypes.byref(d), ctypes.byref(info))
    return a.value, b.value, c.value, d.value, info.value


class OracleMap:
    """A component or triangular map held by the oracle."""

    def __init__(self, handle):
        self.h = ctypes.c_void_p(handle)
        i, o, c = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        lib().orc_info(self.h, ctypes.byref(i), ctypes.byref(o), ctypes.byref(c))
        self.inputDim, self.outputDim, self.numCoeffs = i.value, o.value, c.value

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def SetCoeffs(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.shape == (self.numCoeffs,)
        lib().orc_set_coeffs(self.h, _dp(c))

    def Evaluate(self, pts, return_nevals=False):
        D, N = pts.shape
        out = np.zeros((self.outputDim, N))
        ne = np.zeros((self.outputDim, N), dtype=np.int64) if return_nevals else None
        rs, cs = _strides(pts)
        _check(lib().orc_evaluate(self.h, _dp(pts), D, N, rs, cs, _dp(out), N, 1,
                                  ne.ctypes.data_as(_L) if return_nevals else None))
        return (out, ne) if return_nevals else out

    def LogDeterminant(self, pts):
        D, N = pts.shape
        out = np.zeros(N)
        rs, cs = _strides(pts)
        _check(lib().orc_log_determinant(self.h, _dp(pts), D, N, rs, cs, _dp(out)))
        return out

    def CoeffGrad(self, pts, sens):
        D, N = pts.shape
        sens = np.ascontiguousarray(sens, dtype=np.float64)
        out = np.zeros((self.numCoeffs, N))
        rs, cs = _strides(pts)
        _check(lib().orc_coeff_grad(self.h, _dp(pts), D, N, rs, cs, _dp(sens), N, 1, _dp(out), N, 1))
        return out

    def LogDeterminantCoeffGrad(self, pts):
        D, N = pts.shape
        out = np.zeros((self.numCoeffs, N))
        rs, cs = _strides(pts)
        _check(lib().orc_log_determinant_coeff_grad(self.h, _dp(pts), D, N, rs, cs, _dp(out), N, 1))
        return out

    def Gradient(self, pts, sens):
        D, N = pts.shape
        sens = np.ascontiguousarray(sens, dtype=np.float64)
        out = np.zeros((self.inputDim, N))
        rs, cs = _strides(pts)
        _check(lib().orc_gradient(self.h, _dp(pts), D, N, rs, cs, _dp(sens), N, 1, _dp(out), N, 1))
        return out

    def LogDeterminantInputGrad(self, pts):
        D, N = pts.shape
        out = np.zeros((self.inputDim, N))
        rs, cs = _strides(pts)
        _check(lib().orc_log_determinant_input_grad(self.h, _dp(pts), D, N, rs, cs, _dp(out), N, 1))
        return out

    def Inverse(self, x1, r, xtol=1e-6, ytol=1e-6):
        """Component: x1 is (dim x numXs), r is (1 x N) or (N,).  Triangular: x1 is
        ((inDim-outDim) x N) and r is (outDim x N)."""
        x1 = np.asarray(x1, dtype=np.float64)
        r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
        N = r.shape[1]
        out = np.zeros((self.outputDim, N))
        if x1.size == 0:
            x1 = np.zeros((0, N))
        xrs, xcs = _strides(x1) if x1.ndim == 2 and x1.shape[0] > 0 else (N, 1)
        _check(lib().orc_inverse(self.h, _dp(x1), x1.shape[0], x1.shape[1], xrs, xcs, _dp(r), r.shape[0], N, N, 1,
                                 _dp(out), N, 1, ctypes.c_double(xtol), ctypes.c_double(ytol)))
        return out

    def KL(self, pts, with_const=False):
        D, N = pts.shape
        loss = ctypes.c_double()
        grad = np.zeros(self.numCoeffs)
        rs, cs = _strides(pts)
        _check(lib().orc_kl(self.h, _dp(pts), D, N, rs, cs, int(bool(with_const)), ctypes.byref(loss), _dp(grad)))
        return loss.value, grad


def _new(fn, *args):
    h = ctypes.c_void_p()
    _check(fn(*args, ctypes.byref(h)))
    return OracleMap(h.value)


def CreateComponent(opts, dim, order=None, starts=None, dims=None, orders=None):
    L = lib()
    if starts is None:
        return _new(L.orc_create_total_order_component, ctypes.cast(ctypes.pointer(opts), ctypes.c_void_p), dim, order)
    starts = np.ascontiguousarray(starts, np.uint32)
    dims = np.ascontiguousarray(dims, np.uint32)
    orders = np.ascontiguousarray(orders, np.uint32)
    return _new(L.orc_create_component, ctypes.cast(ctypes.pointer(opts), ctypes.c_void_p), dim, len(starts) - 1, starts.ctypes.data_as(_U),
                dims.ctypes.data_as(_U), orders.ctypes.data_as(_U))


def CreateTriangular(inDim, outDim, totalOrder, opts):
    return _new(lib().orc_create_triangular, ctypes.cast(ctypes.pointer(opts), ctypes.c_void_p), inDim, outDim, totalOrder)


def CreateTriangularOrders(opts, dims, orders):
    dims = np.ascontiguousarray(dims, np.uint32)
    orders = np.ascontiguousarray(orders, np.uint32)
    return _new(lib().orc_create_triangular_orders, ctypes.cast(ctypes.pointer(opts), ctypes.c_void_p), len(dims), dims.ctypes.data_as(_U),
                orders.ctypes.data_as(_U))
