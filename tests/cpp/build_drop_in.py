"""Builds the C++ drop-in test (tests/cpp/drop_in_test.cpp) against
include/intp/Interpolation.hpp and libintp_amd.so.

The known-answer cases of the reference's interpolation test
(test/src/interpolation-test.cpp, restated as data in tests/cases.py from
tests/golden/mma_vectors.json) are written out as C++ blocks that build
each function through the reference-style constructors and evaluate it
through operator(), at and the three derivative_at argument forms; the
explicit-knot cases of test/src/bspline-test.cpp become blocks that build
BSpline<double, D, 3> from knot ranges and control points.  The test
program includes them.

    python tests/cpp/build_drop_in.py [OUT_DIR]   (default tests/cpp/_build)
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

from cases import bspline_cases, interp_cases  # noqa: E402


def _num(x: float) -> str:
    return repr(float(x)) if x == x else "NAN"


def _arr(xs) -> str:
    return "{" + ", ".join(_num(x) for x in xs) + "}"


def _case_block(c) -> str:
    f = c["f"]
    D = f.ndim
    O = c["order"]
    per = [bool(p) for p in c["periodic"]]
    coords = c["coords"] or [None] * D
    ranges = c["ranges"] or [None] * D
    pts = c["points"]
    r = c["orders"][0]
    deriv = any(r)
    lines = ["    {  // %s: %s" % (c["name"], c["cite"])]
    if any(per):
        # without INTP_PERIODIC_NO_DUMMY_POINT a periodic dimension carries
        # one more value, the dummy point, which the build ignores
        # (InterpolationTemplate.hpp:463-486): filled with NaN here
        fd = np.pad(f, [(0, 1) if p else (0, 0) for p in per], constant_values=np.nan)
        lines.append("#ifdef INTP_PERIODIC_NO_DUMMY_POINT")
        lines.append("        static const double f_data[] = %s;" % _arr(f.ravel()))
        lines.append("        const std::array<std::size_t, %d> f_shape{%s};" % (D, ", ".join(str(v) for v in f.shape)))
        lines.append("#else")
        lines.append("        static const double f_data[] = %s;" % _arr(fd.ravel()))
        lines.append("        const std::array<std::size_t, %d> f_shape{%s};" % (D, ", ".join(str(v) for v in fd.shape)))
        lines.append("#endif")
    else:
        lines.append("        static const double f_data[] = %s;" % _arr(f.ravel()))
        lines.append("        const std::array<std::size_t, %d> f_shape{%s};" % (D, ", ".join(str(v) for v in f.shape)))
    for d in range(D):
        if coords[d] is not None:
            lines.append("        static const std::vector<double> nodes%d = %s;" % (d, _arr(coords[d])))
    if D == 1:
        lines.append("        const std::vector<double> fv(f_data, f_data + f_shape[0]);")
        n = f.shape[0]
        if coords[0] is not None:
            xr = "std::make_pair(nodes0.begin(), nodes0.end())"
            lines.append("        const intp::InterpolationFunction1D<%d> fn(%s, intp::util::get_range(fv), %s);"
                         % (O, xr, "true" if per[0] else "false"))
        elif tuple(ranges[0]) == (0.0, float(n if per[0] else n - 1)):
            # default x range of the 1D class
            lines.append("        const intp::InterpolationFunction1D<%d> fn(intp::util::get_range(fv)%s);"
                         % (O, ", true" if per[0] else ""))
        else:
            lines.append("        const intp::InterpolationFunction1D<%d> fn(std::make_pair(%s, %s), "
                         "intp::util::get_range(fv), %s);" % (O, _num(ranges[0][0]), _num(ranges[0][1]),
                                                              "true" if per[0] else "false"))
    else:
        lines.append("        intp::Mesh<double, %d> f(%s);" % (D, ", ".join("f_shape[%d]" % d for d in range(D))))
        lines.append("        std::copy(f_data, f_data + f.size(), f.begin());")
        rs = []
        for d in range(D):
            if coords[d] is not None:
                rs.append("std::make_pair(nodes%d.begin(), nodes%d.end())" % (d, d))
            else:
                rs.append("std::make_pair(%s, %s)" % (_num(ranges[d][0]), _num(ranges[d][1])))
        if any(per):
            lines.append("        const intp::InterpolationFunction<double, %d, %d> fn({%s}, f, %s);"
                         % (D, O, ", ".join("true" if p else "false" for p in per), ", ".join(rs)))
        else:
            lines.append("        const intp::InterpolationFunction<double, %d, %d> fn(f, %s);" % (D, O, ", ".join(rs)))
    lines.append("        static const double pts[%d][%d] = {%s};"
                 % (len(pts), D, ", ".join(_arr(p) for p in pts)))
    lines.append("        static const double want[%d] = %s;" % (len(pts), _arr(c["expected"])))
    lines.append("        std::vector<double> got(%d);" % len(pts))
    lines.append("        for (std::size_t i = 0; i < %d; ++i) {" % len(pts))
    lines.append("            const std::array<double, %d> x{%s};" % (D, ", ".join("pts[i][%d]" % d for d in range(D))))
    if deriv:
        if D == 1:
            call = "fn.derivative_at(std::make_pair(x[0], %d))" % r[0]
        elif D == 2:
            call = "fn.derivative_at(x, %s)" % ", ".join(str(v) for v in r)
        else:
            call = "fn.derivative_at(x, {%s})" % ", ".join(str(v) for v in r)
        if not c["bounds"]:
            call = call.replace("derivative_at", "derivative")
    elif c["bounds"]:
        call = "fn.at(x)"
    else:
        call = "fn(%s)" % ", ".join("x[%d]" % d for d in range(D))
    lines.append("            got[i] = %s;" % call)
    lines.append("        }")
    lines.append('        report("%s", got.data(), want, %d, %r);' % (c["name"], len(pts), c["tol"]))
    # the same points through one batched call
    lines.append("        std::vector<double> batch(%d);" % len(pts))
    lines.append("        const std::size_t r[%d] = {%s};" % (D, ", ".join(str(v) for v in r)))
    lines.append("        fn.eval(%d, &pts[0][0], batch.data(), 1, r, %s);" % (len(pts), "true" if c["bounds"] else "false"))
    lines.append('        report("%s/batched", batch.data(), want, %d, %r);' % (c["name"], len(pts), c["tol"]))
    lines.append("    }")
    return "\n".join(lines)


def _bspline_block(c) -> str:
    """One explicit-knot case: BSpline<double, D, 3> built from knot ranges
    and a control-point mesh (the reference's bspline-test.cpp form),
    evaluated through operator() or derivative_at."""
    ctrl = c["ctrl"]
    D = ctrl.ndim
    O = c["order"]
    per = [bool(p) for p in c["periodic"]]
    pts = c["points"]
    r = c["orders"][0]
    T = "intp::BSpline<double, %d, %d>" % (D, O)
    lines = ["    {  // %s: %s" % (c["name"], c["cite"])]
    for d, k in enumerate(c["knots"]):
        lines.append("        static const std::vector<double> knots%d = %s;" % (d, _arr(k)))
    lines.append("        static const double cp_data[] = %s;" % _arr(ctrl.ravel()))
    lines.append("        %s::ControlPointContainer cp(%s);" % (T, ", ".join(str(s) for s in ctrl.shape)))
    lines.append("        std::copy(cp_data, cp_data + %d, cp.begin());" % ctrl.size)
    ks = ", ".join("std::make_pair(knots%d.begin(), knots%d.end())" % (d, d) for d in range(D))
    if any(per):
        lines.append("        const %s sp({%s}, cp, %s);" % (T, ", ".join("true" if p else "false" for p in per), ks))
    else:
        lines.append("        const %s sp(cp, %s);" % (T, ks))
    lines.append("        static const double pts[%d][%d] = {%s};"
                 % (len(pts), D, ", ".join(_arr(p) for p in pts)))
    lines.append("        static const double want[%d] = %s;" % (len(pts), _arr(c["expected"])))
    lines.append("        std::vector<double> got(%d);" % len(pts))
    if any(r):
        call = "sp.derivative_at({%s})" % ", ".join("std::make_pair(x[%d], %d)" % (d, r[d]) for d in range(D))
        # the (coordinate, hint, order) form, hint = order as the reference passes
        alt = ("sp.derivative_at({%s})"
               % ", ".join("std::make_tuple(x[%d], std::size_t{%d}, std::size_t{%d})" % (d, O, r[d]) for d in range(D)))
        alt_name = "hint tuples"
    else:
        call = "sp(x)"
        # precomputed coefficients applied to the spline (bspline-test.cpp:355-366)
        alt = "sp.pre_calc_coef({%s})(sp)" % ", ".join("std::make_pair(x[%d], std::size_t{%d})" % (d, O) for d in range(D))
        alt_name = "pre_calc_coef"
    lines.append("        std::vector<double> alt(%d);" % len(pts))
    lines.append("        for (std::size_t i = 0; i < %d; ++i) {" % len(pts))
    lines.append("            const std::array<double, %d> x{%s};" % (D, ", ".join("pts[i][%d]" % d for d in range(D))))
    lines.append("            got[i] = %s;" % call)
    lines.append("            alt[i] = %s;" % alt)
    lines.append("        }")
    lines.append('        report("%s", got.data(), want, %d, %r);' % (c["name"], len(pts), c["tol"]))
    lines.append('        report("%s/%s", alt.data(), want, %d, %r);' % (c["name"], alt_name, len(pts), c["tol"]))
    lines.append("    }")
    return "\n".join(lines)


def generate(out_dir: str) -> str:
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "mma_cases.inc")
    body = ["// Generated by tests/cpp/build_drop_in.py from tests/golden/mma_vectors.json."]
    body += [_case_block(c) for c in interp_cases()]
    body += [_bspline_block(c) for c in bspline_cases()]
    text = "\n".join(body) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


def build(out_dir: str = os.path.join(HERE, "_build"), no_dummy: bool = True) -> str:
    """The test program under the reference's test macro set
    (INTP_PERIODIC_NO_DUMMY_POINT, test/CMakeLists.txt:46), or, with
    no_dummy=False, under the reference's default periodic layout
    (drop_in_test_dummy)."""
    generate(out_dir)
    exe = os.path.join(out_dir, "drop_in_test" if no_dummy else "drop_in_test_dummy")
    lib_dir = os.path.join(ROOT, "bsplineinterpolation_amd")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror"]
    cmd += ["-DINTP_PERIODIC_NO_DUMMY_POINT"] if no_dummy else []
    cmd += ["-I" + os.path.join(ROOT, "include"), "-I" + out_dir, os.path.join(HERE, "drop_in_test.cpp"),
            "-o", exe, "-L" + lib_dir, "-lintp_amd", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(lib_dir, out_dir)]
    subprocess.run(cmd, check=True)
    return exe


def build_timing(out_dir: str = os.path.join(HERE, "_build")) -> str:
    """The scalar-call timing program (tests/cpp/drop_in_timing.cpp)."""
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "drop_in_timing")
    lib_dir = os.path.join(ROOT, "bsplineinterpolation_amd")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           "-DINTP_PERIODIC_NO_DUMMY_POINT",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "drop_in_timing.cpp"), "-o", exe,
           "-L" + lib_dir, "-lintp_amd", "-ldl", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(lib_dir, out_dir)]
    subprocess.run(cmd, check=True)
    return exe


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "_build")))
