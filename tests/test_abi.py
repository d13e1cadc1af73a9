"""CPU checks of the product library: it builds, loads, and exports every
entry point include/intp_amd.h declares; host-side argument checks."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "intp_amd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(intp_\w+)\s*\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    from bsplineinterpolation_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from bsplineinterpolation_amd import build
        build.build()
    return _lib.load()


def test_header_declares_expected_entry_points():
    from bsplineinterpolation_amd import _lib
    assert _declared() == sorted(_lib.EXPORTS)


def test_header_constants_match_the_python_binding():
    """Every #define INTP_* number in the header has the same value in
    the ctypes binding (flags, status codes, kinds)."""
    from bsplineinterpolation_amd import _lib
    defs = re.findall(r"^#define\s+(INTP_\w+)\s+(-?\d+)u?\b", open(HEADER).read(), re.M)
    assert len(defs) > 20
    for name, value in defs:
        if hasattr(_lib, name):
            assert getattr(_lib, name) == int(value), name
    for name in ("INTP_EVAL_UNGROUPED", "INTP_EVAL_GROUPED", "INTP_EVAL_OUT_F32", "INTP_BATCH_SPLIT_LINES"):
        assert hasattr(_lib, name), name


def test_library_exports_every_declared_symbol(lib):
    from bsplineinterpolation_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (intp_\w+)", out))
    for name in _declared():
        assert name in exported, name
        assert hasattr(lib, name)


def test_status_strings(lib):
    for code in range(7):
        assert lib.intp_status_string(code) != b"unknown status"
    assert lib.intp_status_string(99) == b"unknown status"
    assert lib.intp_last_error() is not None


def test_serving_entry_points_reject_null_arguments_without_a_device(lib):
    from bsplineinterpolation_amd import _lib
    h = C.c_void_p()
    rid = C.c_int64()
    assert lib.intp_server_create(None, None, C.byref(h)) == _lib.INTP_E_INVALID
    assert lib.intp_server_destroy(None) == _lib.INTP_E_INVALID
    assert lib.intp_serve_build(None, None, None, 0, 0, C.byref(rid)) == _lib.INTP_E_INVALID
    assert lib.intp_serve_eval(None, 1, 1, None, 1, 1, 1, None, None, 1, 1, 0, C.byref(rid)) == _lib.INTP_E_INVALID
    assert lib.intp_serve_wait(None, 1) == _lib.INTP_E_INVALID
    assert lib.intp_serve_drain(None) == _lib.INTP_E_INVALID
    assert lib.intp_serve_stats(None, 0, None) == _lib.INTP_E_INVALID
    st = C.c_int32()
    assert lib.intp_serve_poll(None, 1, C.byref(st), None) == _lib.INTP_E_INVALID


def test_null_arguments_are_rejected_without_a_device(lib):
    from bsplineinterpolation_amd import _lib
    h = C.c_void_p()
    assert lib.intp_template_create(None, None, C.byref(h)) == _lib.INTP_E_INVALID
    assert b"null" in lib.intp_last_error()
    assert lib.intp_eval(None, 1, None, 1, 1, 1, None, None, 1, 1, 0, None) == _lib.INTP_E_INVALID
    assert lib.intp_spline_release(None) == _lib.INTP_OK


def test_product_does_not_import_the_oracle():
    """The product path has no route into the oracle (test infrastructure)."""
    pkg = os.path.join(ROOT, "bsplineinterpolation_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.replace("the oracle", ""), f
                assert "liborc" not in txt, f
