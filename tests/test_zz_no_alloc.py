"""intp_eval_static: evaluation whose scratch (the bounds flag, the grouping
records and tables) comes from the library's fixed per-device buffer
instead of the allocator.  Its outputs must equal intp_eval's bit for bit
(the same kernels), on two calls in a row, the second with a smaller
input (the buffer keeps the first call's contents), grouped and not,
with the bounds check cleared between calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _call(amd, fn, g, pts, orders, flags=0):
    """Evaluates device points [npts, D] (AoS) with intp_eval or
    intp_eval_static into a fresh [npts, nout] output."""
    import torch
    lib = amd._lib.load()
    npts, D = pts.shape
    op = np.ascontiguousarray(np.asarray(orders, np.int32))
    nout = op.shape[0]
    out = torch.full((npts, nout), float("nan"), dtype=torch.float64, device="cuda")
    dp = C.POINTER(C.c_double)
    rc = getattr(lib, fn)(g.h, npts, C.cast(pts.data_ptr(), dp), 1, D, nout, op.ctypes.data_as(C.POINTER(C.c_int32)),
                          C.cast(out.data_ptr(), dp), 1, nout, flags,
                          C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, out


def test_gpu_eval_static_equals_eval(amd):
    import torch
    lib = amd._lib.load()
    cap = lib.intp_eval_static_capacity()
    assert cap >= 64 << 20
    rng = np.random.default_rng(42)
    # a 16.8 MB table: 2^20 points and more are grouped by brick
    f = rng.standard_normal((128, 128, 128))
    g = amd.InterpolationFunction(3, f, periodic=[True, False, False], ranges=[(0, 1), (0, 1), (0, 1)])
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    big = torch.from_numpy(rng.uniform(-0.1, 1.1, ((1 << 20) + 17, 3))).cuda()
    small = torch.from_numpy(rng.uniform(0, 1, (1000, 3))).cuda()
    for pts in (big, small):  # two calls in a row, the second smaller
        rc, got = _call(amd, "intp_eval_static", g, pts, orders)
        assert rc == 0, amd._lib.load().intp_last_error()
        rc, want = _call(amd, "intp_eval", g, pts, orders)
        assert rc == 0
        assert torch.equal(got, want), pts.shape
    # forced grouping of the small batch through the fixed buffer
    rc, got = _call(amd, "intp_eval_static", g, small, orders, amd._lib.INTP_EVAL_GROUPED)
    assert rc == 0 and torch.equal(got, _call(amd, "intp_eval", g, small, orders)[1])
    # the bounds flag lives in the fixed buffer: an out-of-range call,
    # then an in-range call, each reports its own state
    bad = small.clone()
    bad[10, 1] = 1.5
    rc, _ = _call(amd, "intp_eval_static", g, bad, orders[:1], amd._lib.INTP_EVAL_BOUNDS_CHECK)
    assert rc == amd._lib.INTP_E_DOMAIN and b"point 10" in lib.intp_last_error()
    rc, got = _call(amd, "intp_eval_static", g, small, orders[:1], amd._lib.INTP_EVAL_BOUNDS_CHECK)
    assert rc == 0 and torch.equal(got, _call(amd, "intp_eval", g, small, orders[:1])[1])


def test_gpu_eval_static_refined(amd):
    """A refined spline: value and partials come from its gradient table,
    through the fixed buffer as through intp_eval (grouped and not, two
    calls in a row, the second smaller)."""
    import torch
    rng = np.random.default_rng(7)
    f = rng.standard_normal((96, 80, 72))
    g = amd.InterpolationFunction(3, f, periodic=[True, False, True], ranges=[(0, 1)] * 3, precision="f32_refined",
                                  sweeps=1)
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    big = torch.from_numpy(rng.uniform(0, 1, ((1 << 19) + 5, 3))).cuda()
    small = torch.from_numpy(rng.uniform(0, 1, (777, 3))).cuda()
    for flags in (0, amd._lib.INTP_EVAL_GROUPED):
        for pts in (big, small):
            rc, got = _call(amd, "intp_eval_static", g, pts, orders, flags)
            assert rc == 0, amd._lib.load().intp_last_error()
            rc, want = _call(amd, "intp_eval", g, pts, orders, flags)
            assert rc == 0 and torch.equal(got, want), (flags, pts.shape)


def test_gpu_eval_static_limits(amd):
    import torch
    lib = amd._lib.load()
    cap = lib.intp_eval_static_capacity()
    rng = np.random.default_rng(5)
    g = amd.InterpolationFunction(3, rng.standard_normal((64, 64, 64)), ranges=[(0, 1)] * 3)
    # a forced grouping whose records exceed the buffer is refused
    n = cap // 32 + 1024
    pts = torch.rand((n, 3), dtype=torch.float64, device="cuda")
    rc, _ = _call(amd, "intp_eval_static", g, pts, [[0, 0, 0]], amd._lib.INTP_EVAL_GROUPED)
    assert rc == amd._lib.INTP_E_INVALID and b"scratch" in lib.intp_last_error()
    # without forcing, the same batch is evaluated in caller order
    rc, got = _call(amd, "intp_eval_static", g, pts, [[0, 0, 0]])
    assert rc == 0 and torch.equal(got, _call(amd, "intp_eval", g, pts, [[0, 0, 0]], amd._lib.INTP_EVAL_UNGROUPED)[1])
    # host buffers need staging allocations: refused
    q = np.ascontiguousarray(rng.uniform(0, 1, (10, 3)))
    out = np.zeros(10)
    dp = C.POINTER(C.c_double)
    o = np.zeros(3, np.int32)
    rc = lib.intp_eval_static(g.h, 10, q.ctypes.data_as(dp), 1, 3, 1, o.ctypes.data_as(C.POINTER(C.c_int32)),
                              out.ctypes.data_as(dp), 10, 1, amd._lib.INTP_EVAL_HOST_MEM, None)
    assert rc == amd._lib.INTP_E_INVALID
