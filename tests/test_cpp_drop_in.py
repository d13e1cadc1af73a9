"""The C++ drop-in header (include/intp/Interpolation.hpp): the reference's
interpolation known answers and class-surface behaviour, from a compiled
C++ program (tests/cpp/drop_in_test.cpp)."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))

import build_drop_in  # noqa: E402

LIB = os.path.join(os.path.dirname(HERE), "bsplineinterpolation_amd", "libintp_amd.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libintp_amd.so not built")
def test_cpp_drop_in_compiles(tmp_path):
    """Header + generated cases compile warning-free and link (no GPU needed)."""
    exe = build_drop_in.build(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libintp_amd.so not built")
def test_cpp_drop_in_dummy_layout_compiles(tmp_path):
    """The same program without INTP_PERIODIC_NO_DUMMY_POINT (the reference's
    default periodic layout: a dummy point closes each periodic dimension)."""
    exe = build_drop_in.build(str(tmp_path), no_dummy=False)
    assert os.path.exists(exe)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["drop_in_test", "drop_in_test_dummy"])
def test_cpp_drop_in_runs(name):
    exe = os.path.join(HERE, "cpp", "_build", name)
    if not os.path.exists(exe):
        exe = build_drop_in.build(no_dummy=name == "drop_in_test")
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "\n0 of " in "\n" + r.stdout
