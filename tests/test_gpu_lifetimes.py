"""Stream-ordered lifetimes and the evaluation timing hook.

A spline released while its kernels are still queued must stay readable
until they finish (its memory is freed on a release stream ordered after
every recorded use); builds and evaluations on different streams must see
each other's work when the caller orders the streams."""
import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def test_gpu_release_while_queued(amd, orc):
    import torch
    rng = np.random.default_rng(11)
    shape = (120, 96, 80)
    f = rng.standard_normal(shape)
    rg = [(0.0, 1.0), (-1.0, 1.0), (0.0, 2.0)]
    per = [True, False, False]
    fd = torch.from_numpy(f).cuda()
    q = np.stack([rng.uniform(r[0], r[1], 1 << 20) for r in rg])
    qd = torch.from_numpy(q).cuda()
    orders = [[0, 0, 0], [1, 0, 0], [0, 0, 1]]
    side = torch.cuda.Stream()
    outs = []
    with torch.cuda.stream(side):
        for _ in range(4):
            g = amd.InterpolationFunction(3, fd, periodic=per, ranges=rg)
            outs.append(g.eval(qd, orders, stream=side))
            del g  # released with its build and eval still queued
    side.synchronize()
    ref = orc.OracleSpline(3, f, periodic=per, ranges=rg)
    idx = rng.choice(q.shape[1], 2048, replace=False)
    want = ref.eval(np.ascontiguousarray(q[:, idx].T), orders, nthreads=8)
    for o in outs:
        got = o.cpu().numpy()
        assert np.array_equal(got, outs[0].cpu().numpy())
        for i in range(len(orders)):
            assert rel_err(got[i, idx], want[i]) < 1e-12


def test_gpu_builds_sharing_a_grid_on_two_streams(amd, orc):
    """Builds on one uniform grid reuse its uploaded tables; a build issued
    on a second stream right behind the first (which uploaded them) waits
    for that upload, and each spline stays correct after the other and the
    grid's first user are released."""
    import torch
    rng = np.random.default_rng(13)
    shape = (70, 90, 33)
    rg = [(0.0, 1.0), (-2.0, 2.0), (0.5, 1.5)]
    per = [False, True, False]
    fs = [rng.standard_normal(shape) for _ in range(3)]
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    # a fresh range, so the first build uploads the grid
    rg[0] = (0.0, 1.0 + rng.uniform(0.1, 0.2))
    fds = [torch.from_numpy(f).cuda() for f in fs]
    torch.cuda.synchronize()
    with torch.cuda.stream(a):
        g0 = amd.InterpolationFunction(5, fds[0], periodic=per, ranges=rg, stream=a)
    with torch.cuda.stream(b):
        g1 = amd.InterpolationFunction(5, fds[1], periodic=per, ranges=rg, stream=b)
        g2 = amd.InterpolationFunction(5, fds[2], periodic=per, ranges=rg, stream=b, precision="f32_refined")
    q = np.stack([rng.uniform(r[0], r[1], 3000) for r in rg])
    qd = torch.from_numpy(q).cuda()
    b.synchronize()
    del g0
    outs = [g1.eval(qd), g2.eval(qd)]
    torch.cuda.synchronize()
    # refined FP32 on noise data: the refined tests' bar for rough values
    for f, o, bar in ((fs[1], outs[0], 1e-12), (fs[2], outs[1], 1e-5)):
        want = orc.OracleSpline(5, f, periodic=per, ranges=rg).eval(np.ascontiguousarray(q.T), nthreads=8)
        assert rel_err(o.cpu().numpy()[0], want[0]) < bar


def test_gpu_build_and_eval_on_two_streams(amd, orc):
    import torch
    rng = np.random.default_rng(12)
    f = rng.standard_normal((200, 150))
    rg = [(0.0, 3.0), (1.0, 2.0)]
    fd = torch.from_numpy(f).cuda()
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(a):
        g = amd.InterpolationFunction(4, fd, periodic=[False, True], ranges=rg)
    b.wait_stream(a)
    q = np.stack([rng.uniform(r[0], r[1], 5000) for r in rg])
    with torch.cuda.stream(b):
        got = g.eval(torch.from_numpy(q).cuda(), stream=b)
    del g
    b.synchronize()
    want = orc.OracleSpline(4, f, periodic=[False, True], ranges=rg).eval(np.ascontiguousarray(q.T), nthreads=8)
    assert rel_err(got.cpu().numpy()[0], want[0]) < 1e-12
    # control points of a spline from a template, downloaded after release of the template
    t = amd.InterpolationFunctionTemplate(3, (64, 64), ranges=rg)
    s = t.interpolate(fd[:64, :64].contiguous())
    del t
    ctrl = s.control_points()
    o = orc.OracleSpline(3, f[:64, :64].copy(), ranges=rg)
    assert np.array_equal(ctrl, o.control_points())


def test_gpu_eval_timings(amd):
    import torch
    rng = np.random.default_rng(13)
    g = amd.InterpolationFunction(3, rng.standard_normal((64, 64, 64)), ranges=[(0, 1)] * 3)
    assert amd.eval_timings() == []
    small = torch.from_numpy(rng.uniform(0, 1, (3, 1000))).cuda()
    large = torch.from_numpy(rng.uniform(0, 1, (3, 1 << 20))).cuda()
    g.eval(small, timed=True)
    g.eval(large, [[0, 0, 0], [1, 0, 0]], timed=True, group=True)
    g.eval(small)  # not timed
    t = amd.eval_timings()
    assert len(t) == 2
    assert t[0][0] < 0.05 and t[0][1] > 0.0  # small batch: no grouping pass
    assert t[1][0] > 0.0 and t[1][1] > 0.0   # large batch, grouping forced: grouped
    assert amd.eval_timings() == []


def _hip_runtime():
    """The HIP runtime torch loaded (the one this library binds to)."""
    import ctypes
    import glob
    import os

    import torch
    libs = sorted(glob.glob(os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so*")))
    if not libs:
        pytest.skip("torch's HIP runtime library not found")
    return ctypes.CDLL(libs[0])


def test_gpu_stale_runtime_error_not_reported(amd, orc):
    """An error an earlier call left in the HIP runtime's per-thread
    last-error slot is not reported by the next build or evaluation (each
    entry point starts from a clean slot)."""
    import ctypes

    import torch
    hip = _hip_runtime()
    rng = np.random.default_rng(5)
    f = rng.standard_normal((40, 36))
    q = np.stack([rng.uniform(0.0, 1.0, 4096), rng.uniform(0.0, 1.0, 4096)])
    for _ in range(2):
        # hipSetDevice on a device that does not exist fails and changes nothing
        assert hip.hipSetDevice(ctypes.c_int(1 << 20)) != 0
        g = amd.InterpolationFunction(3, torch.from_numpy(f).cuda(), periodic=[False, True],
                                      ranges=[(0.0, 1.0), (0.0, 1.0)])
        assert hip.hipSetDevice(ctypes.c_int(1 << 20)) != 0
        out = g.eval(torch.from_numpy(q).cuda(), [[0, 0], [1, 0]]).cpu().numpy()
    ref = orc.OracleSpline(3, f, periodic=[False, True], ranges=[(0.0, 1.0), (0.0, 1.0)])
    want = ref.eval(np.ascontiguousarray(q.T), [[0, 0], [1, 0]])
    for i in range(2):
        assert rel_err(out[i], want[i]) < 1e-12
