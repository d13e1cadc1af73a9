"""Known-answer cases of the reference test-suite, as data.

Each case names the reference lines it restates.  ``interp_cases()``
covers test/src/interpolation-test.cpp (InterpolationFunction: build +
evaluate, tolerance 1e-14); ``bspline_cases()`` covers
test/src/bspline-test.cpp (explicit knots and control points, 1e-15).
Expected values come from tests/golden/mma_vectors.json or, where the
reference test computes them in closed form (segment / linear), from the
same closed form.
"""
from __future__ import annotations

import bisect
import json
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def load_mma():
    with open(os.path.join(_HERE, "golden", "mma_vectors.json")) as fh:
        return json.load(fh)


def rel_err(got, want) -> float:
    """Norm-wise relative L2 error: test/src/include/rel_err.hpp:11-49."""
    got = np.asarray(got, dtype=np.float64).ravel()
    want = np.asarray(want, dtype=np.float64).ravel()
    return float(math.sqrt(np.sum((got - want) ** 2) / np.sum(want ** 2)))


def interp_cases():
    """List of dicts: name, order, f, periodic, ranges, coords, points,
    orders (one row, the derivative order), expected, tol, bounds."""
    m = load_mma()["interpolation"]
    f1 = np.array(m["f1"])
    tol = 1e-14
    cases = []
    c1 = np.array(m["coords_1d"])
    c1h = np.array(m["coords_1d_half"])

    def add(name, order, f, pts, expected, periodic=None, ranges=None, coords=None, deriv=None,
            bounds=False, cite=""):
        f = np.asarray(f, dtype=np.float64)
        D = f.ndim
        cases.append(dict(name=name, order=order, f=f, periodic=periodic or [False] * D, ranges=ranges,
                          coords=coords, points=np.asarray(pts, dtype=np.float64).reshape(-1, D),
                          orders=[deriv or [0] * D], expected=np.asarray(expected, dtype=np.float64),
                          tol=tol, bounds=bounds, cite=cite))

    # 1D, InterpolationFunction1D<O>(f_range): x range [0, len-1]
    # (Interpolation.hpp:518-533)
    seg = [f1[int(np.floor(x + 0.5))] for x in c1]  # std::round, x > 0
    lin = [f1[int(x)] + (f1[int(x) + 1] - f1[int(x)]) * (x - int(x)) for x in c1]
    add("1d_segment", 0, f1, c1, seg, ranges=[(0.0, 12.0)], cite="interpolation-test.cpp:24,44-54,65")
    add("1d_linear", 1, f1, c1, lin, ranges=[(0.0, 12.0)], cite="interpolation-test.cpp:26,44-54,74")
    add("1d_cubic", 3, f1, c1h, m["vals_1d"], ranges=[(0.0, 6.0)], cite="interpolation-test.cpp:28-29,57-61,83")
    ex = m["extrapolation_1d_cubic"]
    add("1d_cubic_extrapolation", 3, f1, [ex[0][0], ex[1][0]], [ex[0][1], ex[1][1]], ranges=[(0.0, 6.0)],
        cite="interpolation-test.cpp:92-101")
    add("1d_cubic_d1", 3, f1, c1h, m["vals_1d_derivative_1"], ranges=[(0.0, 6.0)], deriv=[1], bounds=True,
        cite="interpolation-test.cpp:433-441")
    # 2D / 3D clamped cubic
    f2 = np.array(m["f2"]).reshape(5, 5)
    c2 = np.array(m["coords_2d"])
    add("2d_cubic", 3, f2, c2, m["vals_2d"], ranges=[(0.0, 4.0), (0.0, 4.0)],
        cite="interpolation-test.cpp:108-151")
    add("2d_cubic_d21", 3, f2, c2, m["vals_2d_derivative_x2_y1"], ranges=[(0.0, 4.0), (0.0, 4.0)],
        deriv=[2, 1], bounds=True, cite="interpolation-test.cpp:472-482")
    f3 = np.array(m["f3"]).reshape(5, 6, 7)
    c3 = np.array(m["coords_3d"])
    r3 = [(0.0, 4.0), (0.0, 5.0), (0.0, 6.0)]
    add("3d_cubic", 3, f3, c3, m["vals_3d"], ranges=r3, cite="interpolation-test.cpp:171-297")
    add("3d_cubic_d103", 3, f3, c3, m["vals_3d_derivative_x1_y0_z3"], ranges=r3, deriv=[1, 0, 3], bounds=True,
        cite="interpolation-test.cpp:494-504")
    # 1D periodic, 12 stored values, x range [0, 12] (NO_DUMMY)
    fp = f1[:12]
    pseg = [f1[int(np.floor(x + 0.5))] for x in c1]
    plin = [f1[int(x)] + (f1[(int(x) + 1) % 12] - f1[int(x)]) * (x - int(x)) for x in c1]
    add("1d_periodic_segment", 0, fp, c1, pseg, periodic=[True], ranges=[(0.0, 12.0)],
        cite="interpolation-test.cpp:313-314,333-352")
    add("1d_periodic_linear", 1, fp, c1, plin, periodic=[True], ranges=[(0.0, 12.0)],
        cite="interpolation-test.cpp:316,333-361")
    add("1d_periodic_quartic", 4, fp, c1, m["vals_1d_periodic"], periodic=[True], ranges=[(0.0, 12.0)],
        cite="interpolation-test.cpp:318,345-370")
    add("1d_periodic_quartic_wrap", 4, fp, [c1[0] - 12.0, c1[0] + 12.0], [m["vals_1d_periodic"][0]] * 2,
        periodic=[True], ranges=[(0.0, 12.0)], cite="interpolation-test.cpp:379-384")
    add("1d_periodic_quartic_d1", 4, fp, c1, m["vals_1d_derivative_periodic"], periodic=[True],
        ranges=[(0.0, 12.0)], deriv=[1], bounds=True, cite="interpolation-test.cpp:449-457")
    # 2D, y periodic (5 x 4 stored)
    add("2d_y_periodic", 3, f2[:, :4], c2, m["vals_2d_periodic"], periodic=[False, True],
        ranges=[(0.0, 4.0), (0.0, 4.0)], cite="interpolation-test.cpp:389-415")
    # non-uniform 1D
    xs = np.array(m["input_coords_1d"])
    nlin = []
    for x in c1:
        i = bisect.bisect_left(list(xs), x)
        nlin.append(f1[i - 1] + (f1[i] - f1[i - 1]) * (x - xs[i - 1]) / (xs[i] - xs[i - 1]))
    add("1d_nonuniform_linear", 1, f1, c1, nlin, coords=[xs], cite="interpolation-test.cpp:517-535,541-559")
    add("1d_nonuniform_cubic", 3, f1, c1, m["vals_1d_nonuniform"], coords=[xs],
        cite="interpolation-test.cpp:537-538,553-568")
    nplin = []
    for x in c1:
        i = bisect.bisect_left(list(xs), x)
        nplin.append(f1[i - 1] + (f1[i % 12] - f1[i - 1]) * (x - xs[i - 1]) / (xs[i] - xs[i - 1]))
    add("1d_nonuniform_periodic_linear", 1, fp, c1, nplin, periodic=[True], coords=[xs],
        cite="interpolation-test.cpp:582-583,596-614")
    add("1d_nonuniform_periodic_quartic", 4, fp, c1, m["vals_1d_nonuniform_periodic"], periodic=[True],
        coords=[xs], cite="interpolation-test.cpp:585-586,608-623")
    # 2D x periodic uniform, y non-uniform (4 x 5 stored)
    add("2d_x_periodic_y_nonuniform", 3, f2[:4, :], c2, m["vals_2d_X_periodic_Y_nonuniform"],
        periodic=[True, False], ranges=[(0.0, 4.0), None], coords=[None, np.array(m["nonuniform_coord_for_2d"])],
        cite="interpolation-test.cpp:636-663")
    return cases


def bspline_cases():
    """Explicit-knot BSpline<double, D, 3> cases (bspline-test.cpp)."""
    b = load_mma()["bspline"]
    k1 = np.array(b["knots"])
    k2 = np.array(b["knots2"])
    cp1 = np.array(b["cp1"])
    cp2 = np.array(b["cp2"], dtype=np.float64).reshape(5, 5)
    cp3 = np.array(b["cp3"], dtype=np.float64).reshape(5, 5, 5)
    c1 = np.array(b["coords_1d"])
    c2 = np.array(b["coords_2d"])
    c3 = np.array(b["coords_3d"])
    out = []

    def add(name, knots, ctrl, pts, r, expected, periodic=None, cite=""):
        D = ctrl.ndim
        out.append(dict(name=name, order=3, knots=knots, ctrl=ctrl, periodic=periodic or [False] * D,
                        points=np.asarray(pts, dtype=np.float64).reshape(-1, D), orders=[r],
                        expected=np.asarray(expected), tol=1e-15, cite=cite))

    add("bs_1d", [k1], cp1, c1, [0], b["vals_1d"], cite="bspline-test.cpp:49-74")
    add("bs_2d", [k1, k1], cp2, c2, [0, 0], b["vals_2d"], cite="bspline-test.cpp:82-123")
    add("bs_3d", [k1, k1, k1], cp3, c3, [0, 0, 0], b["vals_3d"], cite="bspline-test.cpp:131-210")
    add("bs_2d_periodic", [k1, k2], cp2, c2, [0, 0], b["vals_2d_periodic"], periodic=[False, True],
        cite="bspline-test.cpp:219-236")
    add("bs_1d_d0", [k1], cp1, c1, [0], b["vals_1d"], cite="bspline-test.cpp:255-262")
    add("bs_1d_d1", [k1], cp1, c1, [1], b["vals_1d_derivative_1"], cite="bspline-test.cpp:245-249,267-274")
    add("bs_1d_d2", [k1], cp1, c1, [2], b["vals_1d_derivative_2"], cite="bspline-test.cpp:250-254,279-286")
    add("bs_2d_d20", [k1, k1], cp2, c2, [2, 0], b["vals_2d_derivative_x2_y0"], cite="bspline-test.cpp:295-315")
    add("bs_2d_d11", [k1, k1], cp2, c2, [1, 1], b["vals_2d_derivative_x1_y1"], cite="bspline-test.cpp:300-329")
    add("bs_2d_periodic_d11", [k1, k2], cp2, c2, [1, 1], b["vals_2d_periodic_derivative_x1_y1"],
        periodic=[False, True], cite="bspline-test.cpp:336-351")
    return out
