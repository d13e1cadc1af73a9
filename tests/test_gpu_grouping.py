"""Grouped evaluation (large batches bucketed by brick before evaluating)
against the caller-order path and the oracle.

Each point runs the same arithmetic in both paths, so the results must be
bit-identical; a subsample is also checked against the oracle (1e-12).
Grouping is forced (INTP_EVAL_GROUPED): these tables are small enough that
the library would otherwise evaluate them in caller order."""
import math

import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu

UNGROUPED = 4


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _eval_both(g, q, orders, method=True):
    import torch
    qd = torch.from_numpy(q).cuda()
    a = g.eval(qd, orders, group=method).cpu().numpy()
    # same call, caller order
    import ctypes as C
    from bsplineinterpolation_amd import _lib
    nout = len(orders)
    D = g.dim
    out = torch.empty((nout, q.shape[1]), dtype=torch.float64, device="cuda")
    op = np.ascontiguousarray(np.asarray(orders, np.int32))
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().intp_eval(g.h, q.shape[1], C.cast(qd.data_ptr(), dp), q.shape[1], 1, nout,
                                     op.ctypes.data_as(C.POINTER(C.c_int32)), C.cast(out.data_ptr(), dp),
                                     q.shape[1], 1, UNGROUPED, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return a, out.cpu().numpy()


# method True: two radix passes over small bricks (the default);
# "one_pass": one counting pass over large bricks (brick items)
@pytest.mark.parametrize("method", [True, "one_pass"])
@pytest.mark.parametrize("shape,k,per,npts", [
    ((96, 80, 64), 3, [True, False, False], (1 << 21) + 3),
    ((40, 33, 50), 5, [False, True, True], (1 << 20)),
    ((300, 200), 4, [True, True], (1 << 20) + 77),
    ((257, 129), 3, [False, False], (1 << 20)),
    ((17, 9, 300), 2, [True, True, False], (1 << 20) + 1),
    ((1000, 8), 7, [False, True], (1 << 20)),
])
def test_gpu_grouped_equals_ungrouped(amd, orc, shape, k, per, npts, method):
    rng = np.random.default_rng(sum(shape) + k)
    f = rng.standard_normal(shape)
    D = len(shape)
    rg = [(-1.0 + 0.1 * d, 1.5 + 0.2 * d) for d in range(D)]
    g = amd.InterpolationFunction(k, f, periodic=per, ranges=rg)
    span = np.array([r[1] - r[0] for r in rg])
    lo = np.array([r[0] for r in rg])
    q = (lo[:, None] + span[:, None] * rng.uniform(-0.1, 1.1, (D, npts)))
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)] + [[2] * D]
    grouped, plain = _eval_both(g, q, orders, method)
    assert np.array_equal(grouped, plain)
    o = orc.OracleSpline(k, f, periodic=per, ranges=rg)
    idx = rng.choice(npts, 4096, replace=False)
    want = o.eval(np.ascontiguousarray(q[:, idx].T), orders, nthreads=8)
    for i in range(len(orders)):
        assert rel_err(grouped[i, idx], want[i]) < 1e-12


def test_gpu_grouped_host_buffers_and_bounds(amd):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((64, 48, 40))
    g = amd.InterpolationFunction(3, f, ranges=[(0, 1), (0, 1), (0, 1)])
    q = rng.uniform(0, 1, ((1 << 20) + 5, 3))
    host = g.eval(q, [[0, 0, 0], [0, 1, 0]], group=True)
    import torch
    dev = g.eval(torch.from_numpy(np.ascontiguousarray(q.T)).cuda(), [[0, 0, 0], [0, 1, 0]], group=True).cpu().numpy()
    assert np.array_equal(host, dev)
    # the library's own choice (caller order for this 1 MB table) gives the same bits
    assert np.array_equal(host, g.eval(q, [[0, 0, 0], [0, 1, 0]]))
    q[777, 1] = 1.5
    with pytest.raises(amd.DomainError):
        g.eval(q, bounds=True, group=True)
    from bsplineinterpolation_amd import _lib
    import ctypes as C
    qd = torch.from_numpy(np.ascontiguousarray(q.T)).cuda()
    out = torch.empty((1, q.shape[0]), dtype=torch.float64, device="cuda")
    both = _lib.INTP_EVAL_GROUPED | _lib.INTP_EVAL_UNGROUPED
    dp = C.POINTER(C.c_double)
    o = np.zeros(3, np.int32)
    rc = _lib.load().intp_eval(g.h, q.shape[0], C.cast(qd.data_ptr(), dp), q.shape[0], 1, 1,
                               o.ctypes.data_as(C.POINTER(C.c_int32)), C.cast(out.data_ptr(), dp), q.shape[0], 1,
                               both, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.INTP_E_INVALID


def test_gpu_grouped_skewed_points(amd):
    """Most points in one brick (many brick items of one bin, most bins
    empty), the rest spread out, outside the range on clamped dimensions
    and wrapped on the periodic one: grouped results equal caller order."""
    rng = np.random.default_rng(77)
    f = rng.standard_normal((80, 70, 60))
    g = amd.InterpolationFunction(3, f, periodic=[True, False, False], ranges=[(0, 1), (0, 1), (0, 1)])
    n = (1 << 20) + 333
    q = np.empty((3, n))
    hot = n - 50_000
    q[:, :hot] = 0.41 + 0.01 * rng.random((3, hot))
    q[:, hot:] = rng.uniform(-0.3, 1.3, (3, n - hot))
    q[0, hot: hot + 100] += 7.0
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    for method in (True, "one_pass"):
        grouped, plain = _eval_both(g, q, orders, method)
        assert np.array_equal(grouped, plain), method


@pytest.mark.parametrize("npts", [1, 7, 1000, 2049])
def test_gpu_forced_grouping_of_small_batches(amd, npts):
    """INTP_EVAL_GROUPED on batches far below the automatic threshold
    (one partial chunk, a single point) gives the caller-order bits."""
    import torch
    rng = np.random.default_rng(npts)
    f = rng.standard_normal((40, 36, 30))
    g = amd.InterpolationFunction(3, f, periodic=[False, True, False], ranges=[(0, 1), (0, 2), (-1, 1)])
    q = torch.from_numpy(np.stack([rng.uniform(-0.05, 1.05, npts), rng.uniform(0, 2, npts),
                                   rng.uniform(-1, 1, npts)])).cuda()
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    a = g.eval(q, orders, group=True).cpu().numpy()
    b = g.eval(q, orders, group=False).cpu().numpy()
    c = g.eval(q, orders, group="one_pass").cpu().numpy()
    assert np.array_equal(a, b) and np.array_equal(c, b)
