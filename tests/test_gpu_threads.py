"""The C ABI under concurrent host threads (the reference's evaluation
members are const and reentrant, and its construction is thread-safe:
Interpolation.hpp:132-244, DedicatedThreadPool.hpp:45-79).  ctypes releases
the interpreter lock during each library call, so the threads run in the
library at the same time.  Every result must equal the one-thread result
bit for bit:
  * 8 threads evaluating one spline, each on its own stream, small batches
    (caller order) and large ones (grouped by brick, scratch per call);
  * 8 threads building splines of different data on their own streams;
  * 8 threads making host scalar calls on a spline whose host copy is
    downloaded by whichever call comes first."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = 8


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _run(fn, n=THREADS):
    errors, results = [], [None] * n

    def body(i):
        try:
            results[i] = fn(i)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=body, args=(i,)) for i in range(n)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    return results


def test_gpu_concurrent_evaluation_one_spline(amd):
    import torch
    rng = np.random.default_rng(21)
    f = torch.from_numpy(rng.standard_normal((160, 144, 128))).cuda()
    g = amd.InterpolationFunction(3, f, periodic=[True, False, False], ranges=[(0, 1)] * 3)
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    pts = [torch.from_numpy(rng.uniform(0, 1, ((1 << 20) if i % 2 else 5000, 3))).cuda() for i in range(THREADS)]
    want = [g.eval(p, orders).cpu() for p in pts]
    torch.cuda.synchronize()

    def work(i):
        s = torch.cuda.Stream()
        outs = []
        for _ in range(4):
            outs.append(g.eval(pts[i], orders, stream=s))
        s.synchronize()
        return [o.cpu() for o in outs]

    got = _run(work)
    for i in range(THREADS):
        for o in got[i]:
            assert torch.equal(o, want[i]), i


def test_gpu_concurrent_builds(amd):
    import torch
    rng = np.random.default_rng(22)
    data = [rng.standard_normal((64, 56, 48)) for _ in range(THREADS)]
    pts = torch.from_numpy(rng.uniform(0, 1, (20000, 3))).cuda()
    want = []
    for f in data:
        g = amd.InterpolationFunction(5, f, periodic=[False, True, False], ranges=[(0, 1)] * 3)
        want.append((g.control_points(), g.eval(pts, [[0, 0, 0], [0, 1, 0]]).cpu()))

    def work(i):
        s = torch.cuda.Stream()
        fd = torch.from_numpy(data[i]).cuda()
        torch.cuda.synchronize()
        out = []
        for _ in range(3):
            g = amd.InterpolationFunction(5, fd, periodic=[False, True, False], ranges=[(0, 1)] * 3, stream=s)
            v = g.eval(pts, [[0, 0, 0], [0, 1, 0]], stream=s)
            s.synchronize()
            out.append((g.control_points(), v.cpu()))
        return out

    got = _run(work)
    for i in range(THREADS):
        for c, v in got[i]:
            assert np.array_equal(c, want[i][0]) and torch.equal(v, want[i][1]), i


def test_gpu_concurrent_host_calls(amd):
    rng = np.random.default_rng(23)
    f = rng.standard_normal((40, 36, 30))
    q = rng.uniform(0, 1, (3000, 3))
    orders = [[0, 0, 0], [1, 0, 0], [0, 2, 1]]
    want = amd.InterpolationFunction(3, f, periodic=[True, True, False], ranges=[(0, 1)] * 3).eval_host(q, orders)
    for precision in ("f64", "f32_refined"):
        g = amd.InterpolationFunction(3, f, periodic=[True, True, False], ranges=[(0, 1)] * 3, precision=precision)
        # no call before the threads: the first of them downloads the host copy
        got = _run(lambda i: g.eval_host(q[i::2], orders))
        ref = g.eval_host(q, orders)
        for i in range(THREADS):
            assert np.array_equal(got[i], ref[:, i::2]), (precision, i)
        if precision == "f64":
            assert np.array_equal(ref, want)
