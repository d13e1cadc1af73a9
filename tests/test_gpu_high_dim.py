"""Splines of 4 to 6 dimensions (the reference's templates take any
dimension count, Interpolation.hpp:17-18): device build and evaluation,
the host scalar path, evaluation proxies and the refined form, against the
oracle.  The oracle's 4D results equal the compiled reference's, and its
5D and 6D results the products of pinned 1D splines on separable data
(tests/test_oracle.py, tests/test_oracle_vs_ref.py).

Bars as for 1 to 3 dimensions: control points and host-path results bit
for bit; device results within 1e-12 (rel_err); refined values 1e-6.
"""
import itertools

import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _cfg(rng, D, k, per, nonuni, nmax):
    n = [int(rng.integers(max(k + 2, 2 * (k // 2) + 2, 4), nmax + 1)) for _ in range(D)]
    f = rng.standard_normal(n)
    ranges, coords = [], []
    for d in range(D):
        lo = float(rng.uniform(-2, 0))
        hi = lo + float(rng.uniform(0.5, 3))
        ranges.append((lo, hi))
        if nonuni and d % 2 == 1:
            m = n[d] + per[d]
            x = np.sort(rng.uniform(lo, hi, m))
            x[0], x[-1] = lo, hi
            coords.append(x)
        else:
            coords.append(None)
    return f, ranges, (coords if nonuni else None)


CASES = ([(4, k, per, nu) for k in range(0, 8) for per in ([False] * 4, [True, False, True, False], [True] * 4)
          for nu in (False, True) if not (nu and k == 0)]
         + [(5, k, per, nu) for k in (1, 2, 3) for per in ([False] * 5, [True, False, False, True, False])
            for nu in (False, True)]
         + [(6, k, per, False) for k in (1, 3) for per in ([False] * 6, [True, False] * 3)])


def _ids(c):
    return "D%d_k%d_p%s%s" % (c[0], c[1], "".join(str(int(p)) for p in c[2]), "_nu" if c[3] else "")


def _points(rng, o, D, count):
    lo = np.array([o.range(d)[0] for d in range(D)])
    hi = np.array([o.range(d)[1] for d in range(D)])
    q = lo + rng.uniform(-0.2, 1.2, (count, D)) * (hi - lo)
    q[:2] = lo + 0.5 * (hi - lo)
    return q


def _orders(D, k):
    return [[0] * D] + [list(e) for e in np.eye(D, dtype=int)] + [[min(2, k)] + [1] * (D - 1)] + [[k + 1] * D]


@pytest.mark.parametrize("cfg", CASES, ids=[_ids(c) for c in CASES])
def test_gpu_high_dim_matches_oracle(amd, orc, cfg):
    D, k, per, nu = cfg
    rng = np.random.default_rng(hash((D, k, tuple(per), nu, 41)) & 0xFFFFFFFF)
    f, ranges, coords = _cfg(rng, D, k, per, nu, {4: 9, 5: 7, 6: 5}[D])
    print("D=%d k=%d shape %s periodic %s non-uniform %s" % (D, k, f.shape, per, nu), flush=True)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges, coords=coords)
    g = amd.InterpolationFunction(k, f, periodic=per, ranges=ranges, coords=coords)
    for d in range(D):
        assert g.range(d) == o.range(d)
        assert np.array_equal(g.knots(d), o.knots(d))
    assert np.array_equal(g.control_points(), o.control_points())
    q = _points(rng, o, D, 300)
    orders = _orders(D, k)
    want = o.eval(q, orders)
    got = g.eval(q, orders)
    host = g.eval_host(q, orders)
    for i, r in enumerate(orders):
        if np.all(want[i] == 0):
            assert np.all(got[i] == 0) and np.all(host[i] == 0), r
            continue
        assert rel_err(got[i], want[i]) < TOL, (r, rel_err(got[i], want[i]))
        # the host path keeps the reference's operation order
        assert np.array_equal(host[i], want[i]), (r, int(np.sum(host[i] != want[i])))


@pytest.mark.parametrize("D,k", [(4, 3), (5, 2), (6, 1)])
def test_gpu_high_dim_proxies_and_device_points(amd, orc, D, k):
    import torch
    rng = np.random.default_rng(7 * D + k)
    per = [d % 2 == 0 for d in range(D)]
    f, ranges, _ = _cfg(rng, D, k, per, False, {4: 8, 5: 6, 6: 5}[D])
    t = amd.InterpolationFunctionTemplate(k, f.shape, periodic=per, ranges=ranges)
    g = t.interpolate(f)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges)
    q = _points(rng, o, D, 500)
    want = o.eval(q)[0]
    # host proxy: the scalar call's bits
    hp = t.eval_proxy(q)
    assert np.array_equal(hp(g), want)
    # device proxy and device points (AoS and SoA)
    qd = torch.from_numpy(q).cuda()
    dp = t.eval_proxy(qd)
    assert rel_err(dp(g).cpu().numpy(), want) < TOL
    assert rel_err(g.eval(qd)[0].cpu().numpy(), want) < TOL
    assert rel_err(g.eval(qd.t().contiguous())[0].cpu().numpy(), want) < TOL
    # a clamped coordinate outside its range names the point and dimension
    bad = q.copy()
    for d in range(D):
        if not per[d]:
            bad[:, d] = np.clip(bad[:, d], o.range(d)[0], o.range(d)[1])
    bad[17, 1] = o.range(1)[1] + 1.0
    with pytest.raises(amd.DomainError) as e:
        g.eval(torch.from_numpy(bad).cuda(), bounds=True)
    assert "point 17" in str(e.value)


def test_gpu_high_dim_refined(amd, orc):
    """The refined form in 4D (FP32 table, no gradient table): smooth data
    within the refined bars of the oracle."""
    rng = np.random.default_rng(3)
    shape = (14, 12, 10, 9)
    per = [True, False, False, True]
    axes = [np.arange(n) / n if p else np.linspace(0, 1, n) for n, p in zip(shape, per)]
    grid = np.meshgrid(*axes, indexing="ij")
    f = np.ones(shape)
    for d in range(4):
        f = f * np.cos(2 * np.pi * grid[d] + 0.3 * d)
    f += 1e-3 * rng.standard_normal(shape)
    ranges = [(0.0, 1.0)] * 4
    o = orc.OracleSpline(3, f, periodic=per, ranges=ranges)
    g = amd.InterpolationFunction(3, f, periodic=per, ranges=ranges, precision="f32_refined", sweeps=1)
    c = g.control_points()
    assert float(np.max(np.abs(c - o.control_points())) / np.max(np.abs(o.control_points()))) < 1e-6
    q = _points(rng, o, 4, 2000)
    q = np.clip(q, 0.0, 1.0)
    want = o.eval(q)[0]
    got = g.eval(q)[0]
    assert float(np.sqrt(np.mean((got - want) ** 2) / np.mean(want ** 2))) < 1e-6


def test_gpu_4d_larger_grid(amd, orc):
    """A 4D grid of 1.4M control points (40 x 36 x 32 x 30, order 3, dims 0
    and 2 periodic): 2^18 points of value and the four partials on the
    device, a 3000-point subsample against the oracle."""
    import torch
    rng = np.random.default_rng(404)
    shape = (40, 36, 32, 30)
    per = [True, False, True, False]
    ranges = [(0.0, 1.0), (-1.0, 1.0), (0.0, 2.0), (0.5, 1.5)]
    f = rng.standard_normal(shape)
    print("4D grid %s, 2^18 points" % (shape,), flush=True)
    g = amd.InterpolationFunction(3, f, periodic=per, ranges=ranges)
    o = orc.OracleSpline(3, f, periodic=per, ranges=ranges)
    assert np.array_equal(g.control_points(), o.control_points())
    n = 1 << 18
    q = np.stack([rng.uniform(lo, hi, n) for lo, hi in ranges], axis=1)
    orders = [[0] * 4] + [list(e) for e in np.eye(4, dtype=int)]
    got = g.eval(torch.from_numpy(q).cuda(), orders).cpu().numpy()
    idx = rng.choice(n, 3000, replace=False)
    want = o.eval(q[idx], orders)
    for i in range(len(orders)):
        assert rel_err(got[i][idx], want[i]) < TOL, (orders[i], rel_err(got[i][idx], want[i]))
