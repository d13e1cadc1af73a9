"""Host-side evaluation (INTP_EVAL_HOST_MIRROR, the C++ drop-in's scalar
path) and evaluation proxies (intp_proxy_*, the reference's eval_proxy /
pre_calc_coef), both against the oracle at the FP64 bar (1e-12 norm-wise
relative).  Proxy shapes follow the reference's proxy benchmark
(test/src/interpolation-eval-proxy-test.cpp:86-120,145-171): dimensions
1-3, orders 0-5, 2^p mesh points split over the dimensions as
2^((p + d) / D), even dimensions periodic, range [-pi, pi], values the
product of cos(i dt - pi), 2^15 uniform query points sorted by cell.
Device proxies equal intp_eval bit for bit, host proxies equal host
evaluation bit for bit (same weights, same contraction)."""
import ctypes as C
import math

import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _proxy_case(D, p):
    n = [1 << ((p + i) // D) for i in range(D)]
    per = [i % 2 == 0 for i in range(D)]
    f = None
    for i, m in enumerate(n):
        a = np.cos(np.arange(m) * (2 * math.pi / m) - math.pi)
        f = a if f is None else np.multiply.outer(f, a)
    rng = np.random.default_rng(100 * D + p)
    q = rng.uniform(-math.pi, math.pi, (1 << 15, D))
    # sorted by cell, as the reference benchmark sorts its coordinates
    cell = np.zeros(q.shape[0], np.int64)
    for d in range(D):
        cell = cell * n[d] + np.minimum(((q[:, d] + math.pi) / (2 * math.pi) * n[d]).astype(np.int64), n[d] - 1)
    return f, per, q[np.argsort(cell, kind="stable")]


PROXY_CASES = [(D, k, p) for D in (1, 2, 3) for k in range(6) for p in (10, 16)] + [(3, 3, 24), (2, 5, 20)]


@pytest.mark.parametrize("D,k,p", PROXY_CASES, ids=["D%d_k%d_2^%d" % c for c in PROXY_CASES])
def test_gpu_eval_proxy_matches_oracle(amd, orc, D, k, p):
    import torch
    f, per, q = _proxy_case(D, p)
    ranges = [(-math.pi, math.pi)] * D
    t = amd.InterpolationFunctionTemplate(k, f.shape, periodic=per, ranges=ranges)
    g = t.interpolate(f)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges)
    want = o.eval(q)[0]
    # host proxy (the reference's per-point eval_proxy)
    hp = t.eval_proxy(q)
    got_h = hp(g)
    assert rel_err(got_h, want) < TOL, rel_err(got_h, want)
    assert np.array_equal(got_h, g.eval_host(q)[0])
    assert np.array_equal(got_h, want)  # the oracle's (the reference's operator()) bits
    # device proxy
    qd = torch.from_numpy(q).cuda()
    dp = t.eval_proxy(qd)
    got_d = dp(g).cpu().numpy()
    assert rel_err(got_d, want) < TOL, rel_err(got_d, want)
    assert np.array_equal(got_d, g.eval(qd)[0].cpu().numpy())
    # the same proxies on a second function of the grid, and a proxy made
    # from a spline instead of the template
    f2 = 2.0 * f + 1.0
    g2 = t.interpolate(f2)
    want2 = orc.OracleSpline(k, f2, periodic=per, ranges=ranges).eval(q)[0]
    assert rel_err(hp(g2), want2) < TOL
    assert rel_err(dp(g2).cpu().numpy(), want2) < TOL
    assert np.array_equal(g.eval_proxy(qd)(g2).cpu().numpy(), dp(g2).cpu().numpy())


def test_gpu_eval_proxy_rejects_other_grids_and_checks_bounds(amd):
    import torch
    rng = np.random.default_rng(4)
    t = amd.InterpolationFunctionTemplate(3, (12, 10), ranges=[(0, 1), (0, 2)])
    other = amd.InterpolationFunction(3, rng.standard_normal((12, 10)), ranges=[(0, 1), (0, 3)])
    same = amd.InterpolationFunction(3, rng.standard_normal((12, 10)), ranges=[(0, 1), (0, 2)])
    q = rng.uniform(0, 1, (50, 2))
    for pts in (q, torch.from_numpy(q).cuda()):
        px = t.eval_proxy(pts)
        with pytest.raises(amd.IntpError):
            px(other)
        px(same)
    bad = q.copy()
    bad[7, 1] = 2.5
    with pytest.raises(amd.DomainError):
        t.eval_proxy(bad, bounds=True)
    with pytest.raises(amd.DomainError):
        t.eval_proxy(torch.from_numpy(bad).cuda(), bounds=True)
    # refined splines: the device proxy applies the FP32 table
    fr = amd.InterpolationFunction(3, rng.standard_normal((12, 10)), ranges=[(0, 1), (0, 2)],
                                   precision="f32_refined")
    qd = torch.from_numpy(q).cuda()
    assert torch.equal(t.eval_proxy(qd)(fr), fr.eval(qd)[0])


def _rand_cfg(rng, D, k, per, nonuni):
    n = [int(rng.integers(max(k + 2, 2 * (k // 2) + 2, 5), 17)) for _ in range(D)]
    f = rng.standard_normal(n)
    ranges, coords = [], []
    for d in range(D):
        lo = float(rng.uniform(-2, 0))
        hi = lo + float(rng.uniform(0.5, 3))
        ranges.append((lo, hi))
        if nonuni:
            m = n[d] + per[d]
            x = np.sort(rng.uniform(lo, hi, m))
            x[0], x[-1] = lo, hi
            coords.append(x)
        else:
            coords.append(None)
    return f, ranges, (coords if nonuni else None)


HOST_CONFIGS = [(D, k, per, nu) for D in (1, 2, 3) for k in range(0, 8)
                for per in ([False] * D, [True] + [False] * (D - 1), [True] * D) for nu in (False, True)
                if not (nu and k == 0)]


@pytest.mark.parametrize("cfg", HOST_CONFIGS, ids=["D%d_k%d_p%s%s" % (c[0], c[1], "".join(str(int(p)) for p in c[2]),
                                                                       "_nu" if c[3] else "") for c in HOST_CONFIGS])
def test_gpu_host_eval_matches_oracle(amd, orc, cfg):
    """The scalar path: values and derivatives (first, mixed, above the
    order) from the host copy, against the oracle, bit for bit (the host
    path sums in the reference's order); extrapolation and exact knot hits
    included."""
    D, k, per, nu = cfg
    rng = np.random.default_rng(hash((D, k, tuple(per), nu, 7)) & 0xFFFFFFFF)
    f, ranges, coords = _rand_cfg(rng, D, k, per, nu)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges, coords=coords)
    g = amd.InterpolationFunction(k, f, periodic=per, ranges=ranges, coords=coords)
    lo = np.array([o.range(d)[0] for d in range(D)])
    hi = np.array([o.range(d)[1] for d in range(D)])
    q = lo + rng.uniform(-0.2, 1.2, (200, D)) * (hi - lo)
    for d in range(D):
        t = o.knots(d)
        q[:8, d] = t[k: k + 8] if t.size >= k + 8 else q[:8, d]
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)] + [[min(2, k)] + [1] * (D - 1)] + [[k + 1] * D]
    got = g.eval_host(q, orders)
    want = o.eval(q, orders)
    for i, r in enumerate(orders):
        if np.all(want[i] == 0):
            assert np.all(got[i] == 0), r
        else:
            assert rel_err(got[i], want[i]) < TOL, (r, rel_err(got[i], want[i]))
        # the host path sums in the reference's order: the oracle's bits
        assert np.array_equal(got[i], want[i]), (r, int(np.sum(got[i] != want[i])))


def test_gpu_host_eval_bounds_and_refined(amd, orc):
    rng = np.random.default_rng(12)
    f = rng.standard_normal((9, 11))
    g = amd.InterpolationFunction(3, f, periodic=[False, True], ranges=[(0, 1), (0, 2)])
    q = rng.uniform(0, 1, (20, 2))
    q[5, 0] = 1.5  # clamped dimension out of range
    q[3, 1] = 9.0  # periodic: never out of range
    with pytest.raises(amd.DomainError) as e:
        g.eval_host(q, bounds=True)
    assert "point 5" in str(e.value) and "dimension 0" in str(e.value)
    # the device bounds check names the same point
    import torch
    with pytest.raises(amd.DomainError) as e:
        g.eval(torch.from_numpy(q).cuda(), bounds=True)
    assert "point 5" in str(e.value)
    with pytest.raises(amd.DomainError) as e:
        g.eval(q, bounds=True)
    assert "point 5" in str(e.value)
    # refined splines: the host copy holds the FP32 control points
    fr = amd.InterpolationFunction(3, f, periodic=[False, True], ranges=[(0, 1), (0, 2)], precision="f32_refined")
    ok = np.delete(q, 5, axis=0)
    want = orc.OracleSpline(3, f, periodic=[False, True], ranges=[(0, 1), (0, 2)]).eval(ok)[0]
    assert rel_err(fr.eval_host(ok)[0], want) < 1e-6


def test_gpu_host_staged_outputs_keep_gaps(amd):
    """Host buffers with gaps between outputs: only the outputs are
    written (results are staged densely on the device and scattered on
    the host; nothing is copied to the device first)."""
    lib = amd._lib.load()
    rng = np.random.default_rng(3)
    g = amd.InterpolationFunction(3, rng.standard_normal((10, 12)), ranges=[(0, 1), (0, 1)])
    q = np.ascontiguousarray(rng.uniform(0, 1, (64, 2)))
    orders = np.array([[0, 0], [1, 0]], np.int32)
    out = np.full((64, 5), -7.0)  # point i: outputs at columns 0 and 2
    dp = C.POINTER(C.c_double)
    amd._lib.check(lib.intp_eval(g.h, 64, q.ctypes.data_as(dp), 1, 2, 2, orders.ctypes.data_as(C.POINTER(C.c_int32)),
                                 out.ctypes.data_as(dp), 2, 5, amd._lib.INTP_EVAL_HOST_MEM, None))
    want = g.eval(q, orders)
    assert np.array_equal(out[:, 0], want[0]) and np.array_equal(out[:, 2], want[1])
    assert np.all(out[:, [1, 3, 4]] == -7.0)
