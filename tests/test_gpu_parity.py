"""Parity of the HIP path (through the C ABI) with the reference's known
answers and with the CPU oracle.

Tolerances (norm-wise relative L2, the reference's rel_err):
  * MMA vectors: the reference tests' own bars, 1e-14 (interpolation) and
    1e-15 (explicit-knot B-splines);
  * against the oracle: 1e-12 (the FP64 form's parity bar); control points
    are compared bit for bit (the GPU solve keeps the reference's operation
    order without FMA contraction and divides by the pivot).
"""
import itertools
import math

import numpy as np
import pytest

from cases import bspline_cases, interp_cases, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


CASES = interp_cases()
BS = bspline_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_mma_interpolation(amd, case):
    f = amd.InterpolationFunction(case["order"], case["f"], periodic=case["periodic"], ranges=case["ranges"],
                                  coords=case["coords"])
    got = f.eval(case["points"], case["orders"], bounds=case["bounds"])[0]
    assert rel_err(got, case["expected"]) < case["tol"], case["cite"]


@pytest.mark.parametrize("case", BS, ids=[c["name"] for c in BS])
def test_gpu_mma_bspline(amd, case):
    s = amd.InterpolationFunction.from_knots(3, case["knots"], case["ctrl"], periodic=case["periodic"])
    got = s.eval(case["points"], case["orders"])[0]
    assert rel_err(got, case["expected"]) < case["tol"], case["cite"]


def test_gpu_at_out_of_range_raises(amd):
    """interpolation-test.cpp:159-164."""
    c = [x for x in CASES if x["name"] == "2d_cubic"][0]
    f = amd.InterpolationFunction(3, c["f"], ranges=c["ranges"])
    with pytest.raises(amd.DomainError):
        f.at(-1.0, 1.0)
    with pytest.raises(ValueError):
        f.derivative_at([4.5, 1.0], 1, 0)
    assert f.at(0.0, 4.0) == pytest.approx(c["f"][0, 4], abs=1e-14)


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


CONFIGS = [(D, k, list(per), nu) for D in (1, 2, 3) for k in range(0, 8)
           for per in itertools.product([False, True], repeat=D) for nu in (False, True)
           if not (nu and k == 0) and not (D == 3 and per not in ((False,) * 3, (True, False, False), (True,) * 3))]


@pytest.mark.parametrize("cfg", CONFIGS, ids=["D%d_k%d_p%s%s" % (c[0], c[1], "".join(str(int(p)) for p in c[2]),
                                                                  "_nu" if c[3] else "") for c in CONFIGS])
def test_gpu_matches_oracle(amd, orc, cfg):
    D, k, per, nu = cfg
    rng = np.random.default_rng(hash((D, k, tuple(per), nu)) & 0xFFFFFFFF)
    f, ranges, coords = _rand_cfg(rng, D, k, per, nu)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges, coords=coords)
    g = amd.InterpolationFunction(k, f, periodic=per, ranges=ranges, coords=coords)
    for d in range(D):
        assert g.range(d) == o.range(d)
        assert np.array_equal(g.knots(d), o.knots(d))
    # the solve is bit-identical to the reference's arithmetic
    assert np.array_equal(g.control_points(), o.control_points())
    lo = np.array([o.range(d)[0] for d in range(D)])
    hi = np.array([o.range(d)[1] for d in range(D)])
    q = lo + rng.uniform(-0.2, 1.2, (300, D)) * (hi - lo)
    # knots and data sites exactly
    q[:D + 1] = lo + 0.5 * (hi - lo)
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)] + [[min(2, k)] + [1] * (D - 1)] + [[k + 1] * D]
    got = g.eval(q, orders)
    want = o.eval(q, orders)
    for i, r in enumerate(orders):
        if np.all(want[i] == 0):
            assert np.all(got[i] == 0), r
        else:
            assert rel_err(got[i], want[i]) < TOL, (r, rel_err(got[i], want[i]))


def test_gpu_template_reuse(amd, orc):
    """InterpolationFunctionTemplate: factor once, many value meshes
    (InterpolationTemplate.hpp:118-143)."""
    rng = np.random.default_rng(5)
    shape = (12, 10, 9)
    t = amd.InterpolationFunctionTemplate(4, shape, periodic=[True, False, True],
                                          ranges=[(0, 1), (0, 2), (-1, 1)])
    for _ in range(3):
        f = rng.standard_normal(shape)
        g = t.interpolate(f)
        o = orc.OracleSpline(4, f, periodic=[True, False, True], ranges=[(0, 1), (0, 2), (-1, 1)])
        assert np.array_equal(g.control_points(), o.control_points())


def test_gpu_empty_and_tiny_batches(amd):
    f = amd.InterpolationFunction(3, np.arange(25.0).reshape(5, 5), ranges=[(0, 1), (0, 1)])
    assert f.eval(np.zeros((0, 2))).shape == (1, 0)
    assert f(0.5, 0.5) == pytest.approx(12.0, abs=1e-12)


def test_gpu_torch_device_buffers(amd, orc):
    """Device-resident values, SoA and AoS coordinates, on torch's stream."""
    import torch
    rng = np.random.default_rng(8)
    f = rng.standard_normal((20, 18, 16))
    per = [True, False, False]
    rg = [(0.0, 2 * math.pi), (0.0, 1.0), (0.0, 1.0)]
    g = amd.InterpolationFunction(3, torch.from_numpy(f).cuda(), periodic=per, ranges=rg)
    o = orc.OracleSpline(3, f, periodic=per, ranges=rg)
    assert np.array_equal(g.control_points(), o.control_points())
    q = np.stack([rng.uniform(0, 2 * math.pi, 5000), rng.uniform(0, 1, 5000), rng.uniform(0, 1, 5000)])
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    soa = g.eval(torch.from_numpy(q).cuda(), orders).cpu().numpy()
    aos = g.eval(torch.from_numpy(np.ascontiguousarray(q.T)).cuda(), orders).cpu().numpy()
    want = o.eval(q.T, orders)
    assert np.array_equal(soa, aos)
    for i in range(4):
        assert rel_err(soa[i], want[i]) < TOL


def test_gpu_2d_64_cubic_workload(amd, orc):
    """2D order-3 clamped on 64x64 over [0,1]^2, 1e6 queries, value and
    d/dx, against the oracle (the reference's CPU-runnable case)."""
    from bsplineinterpolation_amd import datagen
    w = datagen.workload_2d_64_cubic()
    g = amd.InterpolationFunction(3, w["f"], periodic=w["periodic"], ranges=w["ranges"])
    o = orc.OracleSpline(3, w["f"], periodic=w["periodic"], ranges=w["ranges"])
    assert np.array_equal(g.control_points(), o.control_points())
    q = np.ascontiguousarray(w["queries"].T)
    got = g.eval(q, w["orders"])
    want = o.eval(q, w["orders"], nthreads=16)
    for i in range(2):
        assert rel_err(got[i], want[i]) < TOL
        assert np.max(np.abs(got[i] - want[i])) <= 1e-12 * np.max(np.abs(want[i]))


@pytest.mark.parametrize("k,per", [(2, True), (3, True), (4, False), (5, False), (3, False), (6, True)])
def test_gpu_highest_derivative_on_every_knot(amd, orc, k, per):
    """The order-k derivative is piecewise constant, so evaluating it exactly
    on a knot shows which span was chosen: it must be the reference's
    (t[mu] <= x < t[mu+1]).  Ranges are offset from 0 so that knots are
    not exact multiples of dx (an FMA in the knot formula would move them)."""
    rng = np.random.default_rng(40 + k)
    n = 23
    f = rng.standard_normal(n)
    rg = [(-0.4620982391885941, 0.7015415866960515)]
    o = orc.OracleSpline(k, f, periodic=[per], ranges=rg)
    g = amd.InterpolationFunction(k, f, periodic=[per], ranges=rg)
    t = o.knots(0)
    lo, hi = o.range(0)
    pts = t[(t >= lo) & (t <= hi)][:, None]
    for r in (k, k - 1, 0):
        want = o.eval(pts, [[r]])[0]
        got = g.eval(pts, [[r]])[0]
        assert rel_err(got, want) < TOL, r


@pytest.mark.parametrize("shape,k,per", [
    ((12000,), 3, [True]),        # line longer than an LDS tile: shift copy + in-place HBM solve
    ((12000,), 4, [False]),
    ((8, 10500), 5, [True, True]),
    ((300, 40, 33), 3, [True, False, True]),   # tiles straddling outer lines (inner % lines != 0)
    ((7, 257, 129), 6, [False, True, False]),
])
def test_gpu_solve_tiling_and_long_lines(amd, orc, shape, k, per):
    """Control points bit-identical to the oracle for lines that fit an LDS
    tile (first pass fused with the periodic shift) and lines that do not."""
    rng = np.random.default_rng(sum(shape) + k)
    f = rng.standard_normal(shape)
    rg = [(0.0, 1.0 + d) for d in range(len(shape))]
    o = orc.OracleSpline(k, f, periodic=per, ranges=rg)
    g = amd.InterpolationFunction(k, f, periodic=per, ranges=rg)
    assert np.array_equal(g.control_points(), o.control_points())


def test_gpu_device_points_shape_checked(amd):
    """Device point tensors whose shape does not match the spline's
    dimension raise before any launch."""
    import torch
    rng = np.random.default_rng(8)
    g = amd.InterpolationFunction(3, rng.standard_normal((8, 9, 10)), ranges=[(0, 1)] * 3)
    with pytest.raises(ValueError):
        g.eval(torch.zeros((50, 2), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        g.eval(torch.zeros(3, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        g.eval(torch.zeros((50, 3), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        g.eval(torch.zeros((50, 3), dtype=torch.float64, device="cuda"),
               out=torch.empty((1, 49), dtype=torch.float64, device="cuda"))
    ok = g.eval(torch.full((50, 3), 0.5, dtype=torch.float64, device="cuda"))
    assert ok.shape == (1, 50)
