"""ctypes binding of libintp_amd.so (include/intp_amd.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libintp_amd.so")

MAX_DIM = 6
MAX_ORDER = 7

INTP_OK = 0
INTP_E_INVALID = 1
INTP_E_DOMAIN = 2
INTP_E_NOMEM = 3
INTP_E_HIP = 4
INTP_E_UNSUPPORTED = 5
INTP_E_EVICTED = 6

INTP_MEM_DEVICE = 0
INTP_MEM_HOST = 1
INTP_EVAL_BOUNDS_CHECK = 1
INTP_EVAL_HOST_MEM = 2
INTP_EVAL_UNGROUPED = 4
INTP_EVAL_TIMED = 8
INTP_EVAL_OUT_F32 = 16
INTP_EVAL_GROUPED = 32
INTP_EVAL_HOST_MIRROR = 64
INTP_EVAL_GROUP_ONE_PASS = 128
INTP_BATCH_GRAPH = 1
INTP_BATCH_AUTOTUNE = 2
INTP_BATCH_SPLIT_LINES = 4
INTP_REQ_BUILD = 0
INTP_REQ_EVAL = 1
INTP_REQ_QUEUED = 0
INTP_REQ_RUNNING = 1
INTP_REQ_DONE = 2
INTP_REQ_FAILED = 3
INTP_PREC_F64 = 0
INTP_PREC_F32_REFINED = 1

# every symbol include/intp_amd.h declares
EXPORTS = (
    "intp_last_error", "intp_status_string", "intp_ctx_create", "intp_ctx_destroy",
    "intp_template_create", "intp_template_destroy", "intp_interpolate", "intp_build",
    "intp_spline_from_knots", "intp_spline_retain", "intp_spline_release", "intp_spline_get_info",
    "intp_spline_download_ctrl", "intp_spline_download_knots", "intp_eval", "intp_eval_timings",
    "intp_interpolate_refined", "intp_build_refined",
    "intp_batch_create", "intp_batch_destroy", "intp_batch_info", "intp_batch_build", "intp_batch_eval",
    "intp_batch_run", "intp_batch_tuning", "intp_batch_download_ctrl",
    "intp_server_create", "intp_server_destroy", "intp_serve_build", "intp_serve_eval", "intp_serve_poll",
    "intp_serve_wait", "intp_serve_drain", "intp_serve_release", "intp_serve_spline", "intp_serve_stats",
    "intp_proxy_create", "intp_proxy_eval", "intp_proxy_points", "intp_proxy_destroy",
    "intp_eval_static", "intp_eval_static_capacity",
)


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("order", C.c_int32), ("periodic", C.c_int32 * MAX_DIM),
                ("n", C.c_int64 * MAX_DIM), ("xmin", C.c_double * MAX_DIM), ("xmax", C.c_double * MAX_DIM),
                ("coords", C.POINTER(C.c_double) * MAX_DIM)]


class SplineInfo(C.Structure):
    _fields_ = [("dim", C.c_int32), ("order", C.c_int32), ("periodic", C.c_int32 * MAX_DIM),
                ("uniform", C.c_int32 * MAX_DIM), ("n", C.c_int64 * MAX_DIM), ("nknots", C.c_int64 * MAX_DIM),
                ("range_lo", C.c_double * MAX_DIM), ("range_hi", C.c_double * MAX_DIM),
                ("ctrl_bytes", C.c_int32), ("refine_sweeps", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("streams", C.c_int32), ("max_inflight", C.c_int32)]


class ServeTypeCounters(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("bytes", C.c_double),
                ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_mean_ms", C.c_double), ("latency_max_ms", C.c_double)]


class ServeCounters(C.Structure):
    _fields_ = [("type", ServeTypeCounters * 2), ("resident_splines", C.c_int64), ("resident_bytes", C.c_int64),
                ("evictions", C.c_int64), ("admission_waits", C.c_int64), ("first_submit_s", C.c_double),
                ("last_done_s", C.c_double)]


_LIB = None


def load(path: str = LIB_PATH):
    """Loads libintp_amd.so; raises if it is missing (build it first with
    ``python -m bsplineinterpolation_amd.build``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError("libintp_amd.so is not built (%s); run `python -m bsplineinterpolation_amd.build`" % path)
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # SONAME as ROCm's).  Loading torch first makes this library bind to
    # that copy; loading ours first would leave torch with a second runtime
    # that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, dp, i64p, i32p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    lib.intp_last_error.restype = C.c_char_p
    lib.intp_status_string.restype = C.c_char_p
    lib.intp_status_string.argtypes = [C.c_int]
    lib.intp_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.intp_ctx_destroy.argtypes = [vp]
    lib.intp_template_create.argtypes = [vp, C.POINTER(GridDesc), C.POINTER(vp)]
    lib.intp_template_destroy.argtypes = [vp]
    lib.intp_interpolate.argtypes = [vp, dp, C.c_int, C.POINTER(vp), vp]
    lib.intp_build.argtypes = [vp, C.POINTER(GridDesc), dp, C.c_int, C.POINTER(vp), vp]
    lib.intp_spline_from_knots.argtypes = [vp, C.c_int32, C.c_int32, i32p, i64p, C.POINTER(dp), i64p, dp,
                                           C.POINTER(vp)]
    lib.intp_spline_retain.argtypes = [vp]
    lib.intp_spline_release.argtypes = [vp]
    lib.intp_spline_get_info.argtypes = [vp, C.POINTER(SplineInfo)]
    lib.intp_spline_download_ctrl.argtypes = [vp, dp, C.c_int64]
    lib.intp_spline_download_knots.argtypes = [vp, C.c_int32, dp, C.c_int64]
    lib.intp_eval.argtypes = [vp, C.c_int64, dp, C.c_int64, C.c_int64, C.c_int32, i32p, dp, C.c_int64,
                              C.c_int64, C.c_uint32, vp]
    lib.intp_eval_timings.argtypes = [dp, dp, C.c_int64, i64p]
    lib.intp_interpolate_refined.argtypes = [vp, dp, C.c_int, C.c_int32, C.POINTER(vp), vp]
    lib.intp_build_refined.argtypes = [vp, C.POINTER(GridDesc), dp, C.c_int, C.c_int32, C.POINTER(vp), vp]
    lib.intp_batch_create.argtypes = [vp, C.c_int32, C.POINTER(GridDesc), C.POINTER(vp)]
    lib.intp_batch_destroy.argtypes = [vp]
    lib.intp_batch_info.argtypes = [vp, i64p, i64p]
    lib.intp_batch_build.argtypes = [vp, dp, vp]
    lib.intp_batch_eval.argtypes = [vp, dp, i64p, dp, vp]
    lib.intp_batch_run.argtypes = [vp, dp, dp, i64p, dp, C.c_uint32, vp]
    lib.intp_batch_tuning.argtypes = [vp, C.c_char_p, C.c_int64]
    lib.intp_batch_download_ctrl.argtypes = [vp, C.c_int32, dp, C.c_int64]
    lib.intp_server_create.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(vp)]
    lib.intp_server_destroy.argtypes = [vp]
    lib.intp_serve_build.argtypes = [vp, C.POINTER(GridDesc), dp, C.c_int32, C.c_int32, i64p]
    lib.intp_serve_eval.argtypes = [vp, C.c_int64, C.c_int64, dp, C.c_int64, C.c_int64, C.c_int32, i32p, dp,
                                    C.c_int64, C.c_int64, C.c_uint32, i64p]
    lib.intp_serve_poll.argtypes = [vp, C.c_int64, i32p, i32p]
    lib.intp_serve_wait.argtypes = [vp, C.c_int64]
    lib.intp_serve_drain.argtypes = [vp]
    lib.intp_serve_release.argtypes = [vp, C.c_int64]
    lib.intp_serve_spline.argtypes = [vp, C.c_int64, C.POINTER(vp)]
    lib.intp_serve_stats.argtypes = [vp, C.c_int32, C.POINTER(ServeCounters)]
    lib.intp_proxy_create.argtypes = [vp, vp, C.c_int64, dp, C.c_int64, C.c_int64, C.c_uint32, C.POINTER(vp), vp]
    lib.intp_proxy_eval.argtypes = [vp, vp, dp, C.c_int64, C.c_uint32, vp]
    lib.intp_proxy_points.argtypes = [vp, i64p]
    lib.intp_proxy_destroy.argtypes = [vp]
    lib.intp_eval_static.argtypes = lib.intp_eval.argtypes
    lib.intp_eval_static_capacity.argtypes = []
    lib.intp_eval_static_capacity.restype = C.c_int64
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype or C.c_int
    lib.intp_last_error.restype = C.c_char_p
    lib.intp_status_string.restype = C.c_char_p
    _LIB = lib
    return lib


class IntpError(RuntimeError):
    """A non-zero intp_status; ``code`` is the status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class DomainError(IntpError, ValueError):
    """INTP_E_DOMAIN: the reference's std::domain_error."""


def check(rc: int) -> None:
    if rc == INTP_OK:
        return
    lib = load()
    msg = lib.intp_last_error().decode(errors="replace")
    status = lib.intp_status_string(rc).decode()
    if rc == INTP_E_DOMAIN:
        raise DomainError(rc, msg)
    if rc == INTP_E_NOMEM:
        raise MemoryError(msg)
    raise IntpError(rc, "%s: %s" % (status, msg))
