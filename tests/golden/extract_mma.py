"""Transcribes the reference's known-answer vectors into mma_vectors.json.

The reference tests hold their ground truth (values "pre-computed by MMA",
i.e. Mathematica ``Interpolation[..., Method -> "Spline"]``) as inline C++
arrays.  This script reads those arrays -- data only -- at the cited line
ranges of /root/reference/test/src/*.cpp and writes them, with the inputs,
the evaluation call and the tolerance of each case, as a JSON fixture.
Run it in the build container (the GPU host has no reference checkout):

    python tests/golden/extract_mma.py
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference/test/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mma_vectors.json")

_FLOAT = re.compile(r"-?\d*\.\d+(?:[eE][-+]?\d+)?|-?\d+\.(?:[eE][-+]?\d+)?")
_INT = re.compile(r"-?\d+")


def _lines(fname, a, b):
    with open(os.path.join(REF, fname)) as fh:
        lines = fh.read().split("\n")
    return "\n".join(lines[a - 1:b])


def floats(fname, a, b):
    return [float(s) for s in _FLOAT.findall(_lines(fname, a, b))]


def ints(fname, a, b):
    txt = _lines(fname, a, b)
    txt = txt[txt.index("{"):]  # skip the declaration
    return [int(s) for s in _INT.findall(txt)]


def main():
    it = "interpolation-test.cpp"
    bt = "bspline-test.cpp"
    v = {}
    # ---- interpolation-test.cpp --------------------------------------
    v["f1"] = floats(it, 20, 22)
    assert len(v["f1"]) == 13
    v["coords_1d_half"] = floats(it, 31, 35)
    v["coords_1d"] = floats(it, 38, 42)
    v["vals_1d"] = floats(it, 57, 61)
    v["f2"] = floats(it, 109, 118)
    assert len(v["f2"]) == 25
    c2 = floats(it, 130, 139)
    v["coords_2d"] = [c2[i:i + 2] for i in range(0, 20, 2)]
    v["vals_2d"] = floats(it, 143, 146)
    v["f3"] = floats(it, 172, 261)
    assert len(v["f3"]) == 5 * 6 * 7
    c3 = floats(it, 279, 288)
    v["coords_3d"] = [c3[i:i + 3] for i in range(0, 30, 3)]
    v["vals_3d"] = floats(it, 292, 295)
    v["vals_1d_periodic"] = floats(it, 346, 349)
    v["vals_2d_periodic"] = floats(it, 408, 412)
    v["vals_1d_derivative_1"] = floats(it, 433, 437)
    v["vals_1d_derivative_periodic"] = floats(it, 450, 453)
    v["vals_2d_derivative_x2_y1"] = floats(it, 472, 476)
    v["vals_3d_derivative_x1_y0_z3"] = floats(it, 495, 498)
    v["input_coords_1d"] = floats(it, 517, 529)
    assert len(v["input_coords_1d"]) == 13
    v["vals_1d_nonuniform"] = floats(it, 553, 557)
    v["vals_1d_nonuniform_periodic"] = floats(it, 609, 612)
    v["nonuniform_coord_for_2d"] = floats(it, 636, 637)
    v["vals_2d_X_periodic_Y_nonuniform"] = floats(it, 658, 661)
    v["extrapolation_1d_cubic"] = [[-0.5] + floats(it, 92, 92)[1:], [6.5] + floats(it, 97, 97)[1:]]
    for key in list(v):
        if key.startswith("vals_"):
            assert len(v[key]) == 10, key
    # ---- bspline-test.cpp ---------------------------------------------
    b = {}
    b["knots"] = floats(bt, 49, 49)
    b["cp1"] = [1.0, 16.0 / 3, -16.0 / 3, 4.0 / 3, 3.0]  # bspline-test.cpp:50 (expressions)
    b["coords_1d"] = floats(bt, 57, 61)
    b["vals_1d"] = floats(bt, 63, 67)
    b["cp2"] = ints(bt, 82, 86)
    assert len(b["cp2"]) == 25
    cc = floats(bt, 101, 110)
    b["coords_2d"] = [cc[i:i + 2] for i in range(0, 20, 2)]
    b["vals_2d"] = floats(bt, 114, 117)
    b["cp3"] = ints(bt, 131, 155)
    assert len(b["cp3"]) == 125
    cc = floats(bt, 173, 192)
    b["coords_3d"] = [cc[i:i + 3] for i in range(0, 60, 3)]
    b["vals_3d"] = floats(bt, 196, 202)
    assert len(b["vals_3d"]) == 20
    b["knots2"] = floats(bt, 219, 219)
    assert len(b["knots2"]) == 12
    b["vals_2d_periodic"] = floats(bt, 225, 228)
    b["vals_1d_derivative_1"] = floats(bt, 245, 249)
    b["vals_1d_derivative_2"] = floats(bt, 250, 254)
    b["vals_2d_derivative_x2_y0"] = floats(bt, 295, 299)
    b["vals_2d_derivative_x1_y1"] = floats(bt, 300, 304)
    b["vals_2d_periodic_derivative_x1_y1"] = floats(bt, 336, 340)
    doc = {
        "source": "12ff54e/BSplineInterpolation test/src/interpolation-test.cpp and bspline-test.cpp",
        "macros": ["INTP_PERIODIC_NO_DUMMY_POINT", "INTP_CELL_LAYOUT"],
        "tolerance": {"interpolation": 1e-14, "bspline": 1e-15},
        "metric": "norm-wise relative L2, test/src/include/rel_err.hpp:11-49",
        "interpolation": v,
        "bspline": b,
    }
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
