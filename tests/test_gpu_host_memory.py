"""Host (pageable) memory at the C ABI, as the reference's ConditionalMapBase
methods accept HostSpace views (src/ConditionalMapBase.cpp:7-81): inputs are
staged to the device, outputs copied back, and the results equal the device
path's bitwise.  Strided host views (a transposed numpy array, a strided
output vector) and host / device mixes are covered; the calls go straight
through the C ABI with numpy pointers (the Python layer would copy numpy
arrays itself)."""
import ctypes

import numpy as np
import pytest

import mpart_amd as mt
from mpart_amd.data import Stream

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _hv(a):
    """mpart view of a 2-D float64 numpy array (any strides)."""
    return mt._CView(a.ctypes.data, a.shape[0], a.shape[1], a.strides[0] // 8, a.strides[1] // 8)


def _ok(rc):
    assert rc == 0, mt._lib.mpart_last_error().decode()


def test_host_memory_classification():
    """Which pointers the library stages, checked before any kernel reads one."""
    print("[gpu-case] host_memory classification", flush=True)
    a = np.zeros(16)
    assert mt._lib.mpart_is_host_memory(a.ctypes.data) == 1
    assert mt._lib.mpart_is_host_memory(a[3:].ctypes.data) == 1
    pinned = torch.zeros(16, dtype=torch.float64).pin_memory()
    assert mt._lib.mpart_is_host_memory(pinned.data_ptr()) == 1
    d = torch.zeros(1 << 12, dtype=torch.float64, device="cuda")
    assert mt._lib.mpart_is_host_memory(d.data_ptr()) == 0
    assert mt._lib.mpart_is_host_memory(d[1000:].data_ptr()) == 0
    assert mt._lib.mpart_is_host_memory(None) == 0


def test_host_memory_entry_points():
    print("[gpu-case] host_memory", flush=True)
    a = np.zeros(4)
    assert mt._lib.mpart_is_host_memory(a.ctypes.data) == 1  # before any kernel sees a host pointer
    m = mt.CreateTriangular(3, 3, 3, mt.MapOptions())
    s = Stream(61)
    m.SetCoeffs(0.1 * s.normal(m.numCoeffs))
    N = 777
    xp = s.points(3, N)                       # 3 x N, row-major
    x_host = np.ascontiguousarray(xp.T).T     # the same values, column-major (row stride 1)
    xd = torch.from_numpy(np.ascontiguousarray(xp)).cuda()
    strm = mt._stream()

    # Evaluate: host input (strided) and host output
    y_host = np.zeros((3, N))
    _ok(mt._lib.mpart_evaluate(m._h, _hv(x_host), _hv(y_host), strm))
    y_dev = m.Evaluate(xd).cpu().numpy()
    assert np.array_equal(y_host, y_dev)

    # LogDeterminant into every other entry of a host vector
    ld_host = np.full(2 * N, -7.0)
    _ok(mt._lib.mpart_log_determinant(m._h, _hv(x_host), ld_host.ctypes.data, 2, strm))
    assert np.array_equal(ld_host[0::2], m.LogDeterminant(xd).cpu().numpy())
    assert np.all(ld_host[1::2] == -7.0)

    # CoeffGrad: device points, host sensitivities, host output
    sens = s.normal(3 * N).reshape(3, N)
    cg_host = np.zeros((m.numCoeffs, N))
    _ok(mt._lib.mpart_coeff_grad(m._h, mt._view(xd), _hv(sens), _hv(cg_host), strm))
    assert np.array_equal(cg_host, m.CoeffGrad(xd, torch.from_numpy(sens).cuda()).cpu().numpy())

    # Inverse of T(x): host targets and output
    x1 = np.zeros((0, N))
    xi_host = np.zeros((3, N))
    _ok(mt._lib.mpart_inverse(m._h, _hv(x1), _hv(y_host), _hv(xi_host), 1e-6, 1e-6, strm))
    xi_dev = m.Inverse(torch.zeros((0, N), dtype=torch.float64, device="cuda"),
                       torch.from_numpy(y_host).cuda()).cpu().numpy()
    assert np.array_equal(np.isnan(xi_host), np.isnan(xi_dev))
    assert np.array_equal(np.nan_to_num(xi_host), np.nan_to_num(xi_dev))

    # KL value and gradient into host memory
    loss = np.zeros(1)
    grad = np.zeros(m.numCoeffs)
    _ok(mt._lib.mpart_kl_value_and_grad(m._h, _hv(x_host), ctypes.c_void_p(loss.ctypes.data),
                                        ctypes.c_void_p(grad.ctypes.data), strm))
    ld, gd = m.KLValueAndGrad(xd)
    assert loss[0] == float(ld.item()) and np.array_equal(grad, gd.cpu().numpy())


def test_host_memory_refused_under_capture():
    """Staging cannot live in a graph: host memory during capture is an
    argument error and nothing is launched."""
    print("[gpu-case] host_memory capture", flush=True)
    m = mt.CreateTriangular(2, 2, 2, mt.MapOptions())
    x = np.zeros((2, 8))
    assert mt._lib.mpart_is_host_memory(x.ctypes.data) == 1
    y = torch.zeros((2, 8), dtype=torch.float64, device="cuda")
    z = torch.zeros(4, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
            z.add_(1.0)
            rc = mt._lib.mpart_evaluate(m._h, _hv(x), mt._view(y), ctypes.c_void_p(s.cuda_stream))
    assert rc == -1 and "captured" in mt._lib.mpart_last_error().decode()
    g.replay()
    torch.cuda.synchronize()
    assert torch.all(z == 1.0)
