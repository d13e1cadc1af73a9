"""The C++ drop-in's Mesh / MeshDimension surface (tests/cpp/mesh_test.cpp),
compiled with g++ against include/intp/Interpolation.hpp and run on the
host: indexing, dim_acc_size, iterators along one dimension, resize and
allocators, as the reference's Mesh.hpp defines them."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.parametrize("debug", [False, True])
def test_mesh_surface(tmp_path, debug):
    cxx = os.environ.get("CXX", "g++")
    if shutil.which(cxx) is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path / "mesh_test")
    subprocess.run([cxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include")]
                   + (["-DINTP_DEBUG"] if debug else []) + [os.path.join(HERE, "cpp", "mesh_test.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ("range check" in r.stdout) == debug
