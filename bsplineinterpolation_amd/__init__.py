"""MI355X-native tensor-product B-spline interpolation.

Python mirror of the reference's class surface
(12ff54e/BSplineInterpolation: intp::InterpolationFunction,
InterpolationFunctionTemplate, BSpline) over the C ABI of libintp_amd.so
(include/intp_amd.h).  Every build and evaluation runs the gfx950 kernels;
host (numpy) inputs are staged to the device by the library, device
(torch) tensors are used in place.  The C++ drop-in header
include/intp/Interpolation.hpp offers the same surface to C++ callers.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import DomainError, IntpError, check  # noqa: F401

__all__ = ["Context", "InterpolationFunction", "InterpolationFunctionTemplate", "EvalProxy", "Batch", "Server",
           "DomainError", "IntpError", "eval_timings"]

_dp = C.POINTER(C.c_double)


def _is_torch(x) -> bool:
    return type(x).__module__.split(".")[0] == "torch"


def _stream_handle(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def _current_stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _keep_for_stream(tmp, orig, stream):
    """A temporary made by .contiguous() on the current stream is freed
    when the call returns; if the library reads it on another stream, the
    caching allocator must not hand the block out before that stream is
    done with it."""
    if stream is None or tmp is orig or not hasattr(stream, "cuda_stream"):
        return
    tmp.record_stream(stream)


def _check_device_f64(t, what, numel=None):
    import torch
    if not _is_torch(t) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("%s must be a contiguous float64 GPU tensor" % what)
    if numel is not None and t.numel() != numel:
        raise ValueError("%s must hold %d values, got %d" % (what, numel, t.numel()))


class Context:
    """Binds one GPU (intp_ctx_create)."""

    _defaults = {}

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.intp_ctx_create(int(device), C.byref(h)))
        self.h = h
        self.device = int(device)

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._defaults:
            cls._defaults[device] = cls(device)
        return cls._defaults[device]

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_ctx_destroy(self.h)
            self.h = None


def _grid_desc(order: int, shape: Sequence[int], periodic, ranges, coords):
    D = len(shape)
    if not 1 <= D <= _lib.MAX_DIM:
        raise IntpError(_lib.INTP_E_UNSUPPORTED, "dimension %d has no compiled kernels" % D)
    desc = _lib.GridDesc()
    desc.dim = D
    desc.order = int(order)
    periodic = list(periodic) if periodic is not None else [False] * D
    coords = list(coords) if coords is not None else [None] * D
    ranges = list(ranges) if ranges is not None else [None] * D
    keep = []
    for d in range(D):
        desc.periodic[d] = int(bool(periodic[d]))
        desc.n[d] = int(shape[d])
        if coords[d] is not None:
            c = np.ascontiguousarray(coords[d], dtype=np.float64)
            if c.size != int(shape[d]) + int(bool(periodic[d])):
                raise IntpError(_lib.INTP_E_INVALID,
                                "Inconsistency between knot number and interpolated value number at dimension %d" % d)
            keep.append(c)
            desc.coords[d] = c.ctypes.data_as(_dp)
        else:
            lo, hi = ranges[d]
            desc.xmin[d] = float(lo)
            desc.xmax[d] = float(hi)
    return desc, keep


class InterpolationFunction:
    """intp::InterpolationFunction<double, D, order> (Interpolation.hpp:17-507).

    ``f`` is a C-order array of values (numpy: host, torch: device).
    ``ranges[d]`` = (xmin, xmax) for a uniform dimension, or ``coords[d]`` =
    node coordinates for a non-uniform one (n[d] + periodic[d] of them).
    ``precision``: "f64" (the reference's FP64 arithmetic) or
    "f32_refined" (FP32 control points and evaluation, solved in FP32 with
    ``sweeps`` rounds of FP64-residual refinement, intp_build_refined).
    """

    def __init__(self, order: int, f, periodic: Optional[Sequence[bool]] = None, ranges=None, coords=None,
                 ctx: Optional[Context] = None, stream=None, precision: str = "f64", sweeps: int = 2):
        lib = _lib.load()
        self.ctx = ctx or Context.default()
        desc, keep = _grid_desc(order, tuple(f.shape), periodic, ranges, coords)
        h = C.c_void_p()
        refined = _refined(precision)
        if _is_torch(f):
            f0 = f
            f = f.contiguous()
            if str(f.dtype) != "torch.float64" or not f.is_cuda:
                raise ValueError("device values must be a float64 GPU tensor")
            st = _stream_handle(stream) if stream is not None else _current_stream()
            _keep_for_stream(f, f0, stream)
            vp, mem = C.cast(f.data_ptr(), _dp), _lib.INTP_MEM_DEVICE
        else:
            fa = np.ascontiguousarray(f, dtype=np.float64)
            st = _stream_handle(stream)
            vp, mem = fa.ctypes.data_as(_dp), _lib.INTP_MEM_HOST
        if refined:
            check(lib.intp_build_refined(self.ctx.h, C.byref(desc), vp, mem, int(sweeps), C.byref(h), st))
        else:
            check(lib.intp_build(self.ctx.h, C.byref(desc), vp, mem, C.byref(h), st))
        del keep
        self._set_handle(h)

    @classmethod
    def _wrap(cls, h, ctx):
        self = cls.__new__(cls)
        self.ctx = ctx
        self._set_handle(h)
        return self

    def _set_handle(self, h):
        self.h = h
        info = _lib.SplineInfo()
        check(_lib.load().intp_spline_get_info(h, C.byref(info)))
        self.dim = info.dim
        self.order = info.order
        self._info = info
        self.precision = "f32_refined" if info.ctrl_bytes == 4 else "f64"
        self.refine_sweeps = info.refine_sweeps if info.ctrl_bytes == 4 else None

    @classmethod
    def from_knots(cls, order: int, knots: Sequence[np.ndarray], ctrl: np.ndarray, periodic=None,
                   ctx: Optional[Context] = None) -> "InterpolationFunction":
        """BSpline<double, D, order>(periodicity, ctrl, knot ranges...)."""
        lib = _lib.load()
        ctx = ctx or Context.default()
        ctrl = np.ascontiguousarray(ctrl, dtype=np.float64)
        D = ctrl.ndim
        per = (C.c_int32 * _lib.MAX_DIM)(*[int(bool(p)) for p in (periodic or [False] * D)])
        n = (C.c_int64 * _lib.MAX_DIM)(*ctrl.shape)
        ks = [np.ascontiguousarray(k, dtype=np.float64) for k in knots]
        kp = (_dp * _lib.MAX_DIM)(*[k.ctypes.data_as(_dp) for k in ks])
        nk = (C.c_int64 * _lib.MAX_DIM)(*[k.size for k in ks])
        h = C.c_void_p()
        check(lib.intp_spline_from_knots(ctx.h, D, int(order), per, n, kp, nk, ctrl.ctypes.data_as(_dp),
                                         C.byref(h)))
        return cls._wrap(h, ctx)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_spline_release(self.h)
            self.h = None

    # ---- properties (Interpolation.hpp:248-267) ----------------------
    def periodicity(self, d: int) -> bool:
        return bool(self._info.periodic[d])

    def uniform(self, d: int) -> bool:
        return bool(self._info.uniform[d])

    def range(self, d: int):
        return (self._info.range_lo[d], self._info.range_hi[d])

    def get_order(self) -> int:
        return self.order

    def control_points(self) -> np.ndarray:
        shape = tuple(self._info.n[d] for d in range(self.dim))
        out = np.empty(shape)
        check(_lib.load().intp_spline_download_ctrl(self.h, out.ctypes.data_as(_dp), out.size))
        return out

    def knots(self, d: int) -> np.ndarray:
        out = np.empty(self._info.nknots[d])
        check(_lib.load().intp_spline_download_knots(self.h, d, out.ctypes.data_as(_dp), out.size))
        return out

    # ---- evaluation ---------------------------------------------------
    def eval_host(self, points, orders=None, bounds: bool = False):
        """Evaluation on this thread from a host copy of the control points
        (INTP_EVAL_HOST_MIRROR; the copy is made on the first call): the
        per-point path of the C++ drop-in's scalar members.  ``points``:
        numpy [npts, D]; returns numpy [nout, npts] float64."""
        D = self.dim
        orders = np.zeros((1, D), np.int32) if orders is None else \
            np.ascontiguousarray(np.asarray(orders, dtype=np.int32).reshape(-1, D))
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, D))
        res = np.empty((orders.shape[0], pts.shape[0]))
        flags = _lib.INTP_EVAL_HOST_MIRROR | (_lib.INTP_EVAL_BOUNDS_CHECK if bounds else 0)
        check(_lib.load().intp_eval(self.h, pts.shape[0], pts.ctypes.data_as(_dp), 1, D, orders.shape[0],
                                    orders.ctypes.data_as(C.POINTER(C.c_int32)), res.ctypes.data_as(_dp),
                                    pts.shape[0], 1, flags, None))
        return res

    def eval_proxy(self, points, bounds: bool = False, stream=None) -> "EvalProxy":
        """eval_proxy (Interpolation.hpp:493-506) for many points: spans and
        basis values of ``points`` on this spline's grid, computed once;
        the returned proxy evaluates any spline on an equal grid."""
        return EvalProxy(points, spline=self, bounds=bounds, stream=stream)

    def eval(self, points, orders=None, bounds: bool = False, out=None, stream=None, timed: bool = False,
             out_dtype=None, group: Optional[bool] = None):
        """Batched evaluation.  ``points``: numpy [npts, D] (host) or a
        float64 torch tensor [D, npts] (SoA) / [npts, D] (AoS) on the
        device.  ``orders``: [nout, D] derivative orders (default: value).
        Returns [nout, npts] (numpy or torch, like ``points``), float64, or
        float32 when ``out`` is float32 or ``out_dtype`` says so.  ``timed``
        records device events around the call (read with eval_timings()).
        ``group``: True / False forces / forbids grouping the points by
        brick (INTP_EVAL_GROUPED / INTP_EVAL_UNGROUPED); "one_pass" forces
        the one-pass grouping (INTP_EVAL_GROUP_ONE_PASS); None lets
        the library decide from the batch and table sizes.  Results are the
        same either way."""
        lib = _lib.load()
        D = self.dim
        orders = np.zeros((1, D), np.int32) if orders is None else \
            np.ascontiguousarray(np.asarray(orders, dtype=np.int32).reshape(-1, D))
        nout = orders.shape[0]
        flags = (_lib.INTP_EVAL_BOUNDS_CHECK if bounds else 0) | (_lib.INTP_EVAL_TIMED if timed else 0)
        if group == "one_pass":
            flags |= _lib.INTP_EVAL_GROUPED | _lib.INTP_EVAL_GROUP_ONE_PASS
        elif group is not None:
            flags |= _lib.INTP_EVAL_GROUPED if group else _lib.INTP_EVAL_UNGROUPED
        op = orders.ctypes.data_as(C.POINTER(C.c_int32))
        if _is_torch(points):
            import torch
            if points.dtype != torch.float64 or not points.is_cuda:
                raise ValueError("device points must be a float64 GPU tensor")
            pts = points.contiguous()
            if pts.dim() == 1 and D == 1:
                pts = pts.reshape(1, -1)
            if pts.dim() != 2 or D not in tuple(pts.shape):
                raise ValueError("points must be [%d, npts] (SoA) or [npts, %d] (AoS), got %s"
                                 % (D, D, tuple(pts.shape)))
            if pts.shape[0] == D and (pts.shape[1] != D or D == 1):
                npts, cs_dim, cs_pt = pts.shape[1], pts.shape[1], 1
            else:
                npts, cs_dim, cs_pt = pts.shape[0], 1, D
            if out is None:
                dt = torch.float32 if out_dtype in (torch.float32, np.float32, "float32") else torch.float64
                out = torch.empty((nout, npts), dtype=dt, device=pts.device)
            if (tuple(out.shape) != (nout, npts) or out.dtype not in (torch.float64, torch.float32)
                    or not out.is_cuda or out.device != pts.device):
                raise ValueError("out must be a float64/float32 tensor [%d, %d] on the points' device"
                                 % (nout, npts))
            if out.dtype == torch.float32:
                flags |= _lib.INTP_EVAL_OUT_F32
            st = _stream_handle(stream) if stream is not None else _current_stream()
            _keep_for_stream(pts, points, stream)
            check(lib.intp_eval(self.h, npts, C.cast(pts.data_ptr(), _dp), cs_dim, cs_pt, nout, op,
                                C.cast(out.data_ptr(), _dp), out.stride(0), out.stride(1), flags, st))
            return out
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, D))
        npts = pts.shape[0]
        f32 = out_dtype in (np.float32, "float32")
        res = np.empty((nout, npts), dtype=np.float32 if f32 else np.float64)
        if f32:
            flags |= _lib.INTP_EVAL_OUT_F32
        check(lib.intp_eval(self.h, npts, pts.ctypes.data_as(_dp), 1, D, nout, op,
                            C.cast(res.ctypes.data, _dp), npts, 1, flags | _lib.INTP_EVAL_HOST_MEM,
                            _stream_handle(stream)))
        return res

    def __call__(self, *x):
        """operator()(x...) / operator()(coord array): value, no bounds check."""
        pts = np.asarray(x[0] if len(x) == 1 else x, dtype=np.float64)
        if pts.ndim <= 1:
            return float(self.eval(pts.reshape(1, self.dim))[0, 0])
        return self.eval(pts)[0]

    def at(self, *x):
        """at(x...): value with the clamped-range check (std::domain_error)."""
        pts = np.asarray(x[0] if len(x) == 1 else x, dtype=np.float64).reshape(-1, self.dim)
        v = self.eval(pts, bounds=True)[0]
        return float(v[0]) if v.size == 1 else v

    def derivative(self, coord, *orders):
        """derivative(coord, orders...) (Interpolation.hpp:177-205)."""
        r = np.asarray(orders[0] if len(orders) == 1 else orders, dtype=np.int32).reshape(1, self.dim)
        pts = np.asarray(coord, dtype=np.float64).reshape(-1, self.dim)
        v = self.eval(pts, r)[0]
        return float(v[0]) if v.size == 1 else v

    def derivative_at(self, coord, *orders):
        """derivative_at(coord, orders...): with the range check."""
        r = np.asarray(orders[0] if len(orders) == 1 else orders, dtype=np.int32).reshape(1, self.dim)
        pts = np.asarray(coord, dtype=np.float64).reshape(-1, self.dim)
        v = self.eval(pts, r, bounds=True)[0]
        return float(v[0]) if v.size == 1 else v


def _refined(precision: str) -> bool:
    if precision not in ("f64", "f32_refined"):
        raise IntpError(_lib.INTP_E_INVALID, "precision must be 'f64' or 'f32_refined'")
    return precision == "f32_refined"


def eval_timings(capacity: int = 1 << 16):
    """[(group_ms, kernel_ms)] of the evaluations made with ``timed=True``
    on this thread since the last call (intp_eval_timings), oldest first;
    waits for the last of them."""
    lib = _lib.load()
    g = np.zeros(capacity)
    k = np.zeros(capacity)
    n = C.c_int64()
    check(lib.intp_eval_timings(g.ctypes.data_as(_dp), k.ctypes.data_as(_dp), capacity, C.byref(n)))
    m = min(n.value, capacity)
    return list(zip(g[:m].tolist(), k[:m].tolist()))


class InterpolationFunctionTemplate:
    """intp::InterpolationFunctionTemplate<double, D, order>
    (InterpolationTemplate.hpp:32-581): knots and factored collocation
    matrices built once, then ``interpolate(values)`` per value mesh."""

    def __init__(self, order: int, shape: Sequence[int], periodic=None, ranges=None, coords=None,
                 ctx: Optional[Context] = None):
        lib = _lib.load()
        self.ctx = ctx or Context.default()
        self.shape = tuple(int(s) for s in shape)
        desc, keep = _grid_desc(order, self.shape, periodic, ranges, coords)
        h = C.c_void_p()
        check(lib.intp_template_create(self.ctx.h, C.byref(desc), C.byref(h)))
        del keep
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_template_destroy(self.h)
            self.h = None

    def interpolate(self, f, stream=None, precision: str = "f64", sweeps: int = 2) -> InterpolationFunction:
        lib = _lib.load()
        if tuple(f.shape) != self.shape:
            raise IntpError(_lib.INTP_E_INVALID, "value mesh shape does not match the template")
        h = C.c_void_p()
        if _is_torch(f):
            f0 = f
            f = f.contiguous()
            if str(f.dtype) != "torch.float64" or not f.is_cuda:
                raise ValueError("device values must be a float64 GPU tensor")
            st = _stream_handle(stream) if stream is not None else _current_stream()
            _keep_for_stream(f, f0, stream)
            vp, mem = C.cast(f.data_ptr(), _dp), _lib.INTP_MEM_DEVICE
        else:
            fa = np.ascontiguousarray(f, dtype=np.float64)
            st = _stream_handle(stream)
            vp, mem = fa.ctypes.data_as(_dp), _lib.INTP_MEM_HOST
        if _refined(precision):
            check(lib.intp_interpolate_refined(self.h, vp, mem, int(sweeps), C.byref(h), st))
        else:
            check(lib.intp_interpolate(self.h, vp, mem, C.byref(h), st))
        return InterpolationFunction._wrap(h, self.ctx)

    def eval_proxy(self, points, bounds: bool = False, stream=None) -> EvalProxy:
        """eval_proxy (InterpolationTemplate.hpp:145-176) for many points."""
        return EvalProxy(points, template=self, bounds=bounds, stream=stream)


class EvalProxy:
    """Evaluation proxy (intp_proxy_*; the reference's eval_proxy /
    pre_calc_coef, Interpolation.hpp:493-506, BSpline.hpp:244-297): the knot
    spans and value basis of fixed ``points`` on a grid (of ``template`` or
    of ``spline``), computed once.  ``proxy(f)`` returns the values of any
    spline ``f`` on an equal grid at those points.  numpy points make a
    host proxy (applied on this thread to the spline's host copy of its
    control points); a float64 torch tensor [npts, D] makes a device proxy
    whose values equal ``f.eval`` bit for bit."""

    def __init__(self, points, template=None, spline=None, bounds: bool = False, stream=None):
        lib = _lib.load()
        if (template is None) == (spline is None):
            raise ValueError("give exactly one of template and spline")
        owner = template if template is not None else spline
        D = owner.dim if spline is not None else len(owner.shape)
        self._owner = owner
        flags = _lib.INTP_EVAL_BOUNDS_CHECK if bounds else 0
        self.host = not _is_torch(points)
        if self.host:
            pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, D))
            ptr, st = pts.ctypes.data_as(_dp), None
            flags |= _lib.INTP_EVAL_HOST_MEM
        else:
            import torch
            if points.dtype != torch.float64 or not points.is_cuda:
                raise ValueError("device points must be a float64 GPU tensor")
            pts = points.contiguous().reshape(-1, D)
            ptr = C.cast(pts.data_ptr(), _dp)
            st = _stream_handle(stream) if stream is not None else _current_stream()
        self.npts = int(pts.shape[0])
        self.device = None if self.host else pts.device
        h = C.c_void_p()
        check(lib.intp_proxy_create(template.h if template is not None else None,
                                    spline.h if spline is not None else None, self.npts, ptr, 1, D, flags,
                                    C.byref(h), st))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_proxy_destroy(self.h)
            self.h = None

    def __call__(self, f: "InterpolationFunction", out=None, stream=None):
        lib = _lib.load()
        if self.host:
            res = np.empty(self.npts)
            check(lib.intp_proxy_eval(self.h, f.h, res.ctypes.data_as(_dp), 1, 0, None))
            return res
        import torch
        if out is None:
            out = torch.empty(self.npts, dtype=torch.float64, device=self.device)
        if out.numel() != self.npts or not out.is_cuda or out.dtype not in (torch.float64, torch.float32):
            raise ValueError("out must be a float64/float32 GPU tensor of %d values" % self.npts)
        flags = _lib.INTP_EVAL_OUT_F32 if out.dtype == torch.float32 else 0
        st = _stream_handle(stream) if stream is not None else _current_stream()
        check(lib.intp_proxy_eval(self.h, f.h, C.cast(out.data_ptr(), _dp), out.stride(0) if out.dim() else 1,
                                  flags, st))
        return out


class Batch:
    """Many independent FP64 splines built and evaluated together
    (intp_batch_*).  ``specs``: one dict per spline with ``order``,
    ``shape`` and optionally ``periodic``, ``ranges``, ``coords`` (as for
    InterpolationFunction).  Values are the splines' C-order meshes
    concatenated (``offsets`` gives each start); points are [npts, 3]
    (x, y, z; unused dims ignored), spline s owning
    ``point_offsets[s]:point_offsets[s+1]``."""

    def __init__(self, specs, ctx: Optional[Context] = None):
        lib = _lib.load()
        self.ctx = ctx or Context.default()
        self.count = len(specs)
        descs = (_lib.GridDesc * self.count)()
        keep = []
        for i, sp in enumerate(specs):
            d, k = _grid_desc(sp["order"], tuple(sp["shape"]), sp.get("periodic"), sp.get("ranges"), sp.get("coords"))
            descs[i] = d
            keep.append(k)
        h = C.c_void_p()
        check(lib.intp_batch_create(self.ctx.h, self.count, descs, C.byref(h)))
        self.h = h
        del keep
        total = C.c_int64()
        self.offsets = np.zeros(self.count + 1, np.int64)
        check(lib.intp_batch_info(h, C.byref(total), self.offsets.ctypes.data_as(C.POINTER(C.c_int64))))
        self.total = int(total.value)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_batch_destroy(self.h)
            self.h = None

    def _offsets(self, point_offsets):
        po = np.ascontiguousarray(np.asarray(point_offsets, dtype=np.int64))
        if po.ndim != 1 or po.size != self.count + 1:
            raise ValueError("point_offsets must hold count + 1 = %d entries" % (self.count + 1))
        if po[0] != 0 or np.any(np.diff(po) < 0):
            raise ValueError("point_offsets must start at 0 and be non-decreasing")
        return po, po.ctypes.data_as(C.POINTER(C.c_int64))

    def _points(self, coords, po, out):
        import torch
        _check_device_f64(coords, "coords")
        if coords.dim() != 2 or coords.shape[1] != 3 or coords.shape[0] < int(po[-1]):
            raise ValueError("coords must be [npts, 3] with npts >= point_offsets[-1] = %d" % int(po[-1]))
        if out is None:
            out = torch.empty(int(po[-1]), dtype=torch.float64, device=coords.device)
        _check_device_f64(out, "out")
        if out.numel() < int(po[-1]):
            raise ValueError("out must hold point_offsets[-1] = %d values" % int(po[-1]))
        return out

    def build(self, values, stream=None):
        _check_device_f64(values, "values", self.total)
        st = _stream_handle(stream) if stream is not None else _current_stream()
        check(_lib.load().intp_batch_build(self.h, C.cast(values.data_ptr(), _dp), st))

    def eval(self, coords, point_offsets, out=None, stream=None):
        po, pp = self._offsets(point_offsets)
        out = self._points(coords, po, out)
        st = _stream_handle(stream) if stream is not None else _current_stream()
        check(_lib.load().intp_batch_eval(self.h, C.cast(coords.data_ptr(), _dp), pp, C.cast(out.data_ptr(), _dp), st))
        return out

    def run(self, values, coords, point_offsets, out=None, graph: bool = False, autotune: bool = False,
            split_lines: bool = False, stream=None):
        """Build + eval (device tensors); ``graph`` replays one captured
        hipGraph while buffers and layout stay the same; ``split_lines``
        solves long 1D lines in parallel segments (INTP_BATCH_SPLIT_LINES:
        within 1e-12 of the sequential solve, not bit-identical)."""
        _check_device_f64(values, "values", self.total)
        po, pp = self._offsets(point_offsets)
        out = self._points(coords, po, out)
        flags = ((_lib.INTP_BATCH_GRAPH if graph else 0) | (_lib.INTP_BATCH_AUTOTUNE if autotune else 0) |
                 (_lib.INTP_BATCH_SPLIT_LINES if split_lines else 0))
        st = _stream_handle(stream) if stream is not None else _current_stream()
        check(_lib.load().intp_batch_run(self.h, C.cast(values.data_ptr(), _dp), C.cast(coords.data_ptr(), _dp), pp,
                                         C.cast(out.data_ptr(), _dp), flags, st))
        return out

    def tuning(self) -> str:
        buf = C.create_string_buffer(4096)
        check(_lib.load().intp_batch_tuning(self.h, buf, 4096))
        return buf.value.decode()

    def control_points(self, s: int) -> np.ndarray:
        n = int(self.offsets[s + 1] - self.offsets[s])
        out = np.empty(n)
        check(_lib.load().intp_batch_download_ctrl(self.h, int(s), out.ctypes.data_as(_dp), n))
        return out


class Server:
    """The serving loop (intp_server_* / intp_serve_*): builds and
    evaluations are queued from any thread and run by the library's
    dispatcher thread on ``streams`` HIP streams; builds are admitted
    against ``hbm_budget`` bytes of resident control points, evicting
    least-recently-used idle splines.  Request ids returned by ``build``
    also name the spline for ``eval``.  Tensors passed to a request must
    stay alive until it completes (the server keeps a reference until its
    ``wait``/``drain``)."""

    def __init__(self, ctx: Optional[Context] = None, hbm_budget: int = 0, streams: int = 4,
                 max_inflight: int = 256):
        lib = _lib.load()
        self.ctx = ctx or Context.default()
        cfg = _lib.ServerConfig(int(hbm_budget), int(streams), int(max_inflight))
        h = C.c_void_p()
        check(lib.intp_server_create(self.ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._hold = {}

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None and _lib._LIB is not None:
            _lib._LIB.intp_server_destroy(self.h)
            self.h = None
            self._hold.clear()

    def __del__(self):
        self.close()

    def build(self, order: int, values, periodic=None, ranges=None, coords=None, precision: str = "f64",
              sweeps: int = 2) -> int:
        """Queues a build of device tensor ``values`` (float64, C-order);
        returns the request (= spline) id."""
        desc, keep = _grid_desc(order, tuple(values.shape), periodic, ranges, coords)
        assert _is_torch(values) and str(values.dtype) == "torch.float64", "values must be a float64 device tensor"
        values = values.contiguous()
        prec = _lib.INTP_PREC_F32_REFINED if _refined(precision) else _lib.INTP_PREC_F64
        rid = C.c_int64()
        check(_lib.load().intp_serve_build(self.h, C.byref(desc), C.cast(values.data_ptr(), _dp), prec, int(sweeps),
                                           C.byref(rid)))
        del keep
        self._hold[rid.value] = values
        return rid.value

    def eval(self, spline: int, coords, orders=None, out=None, out_dtype=None) -> int:
        """Queues an evaluation of ``coords`` ([npts, D] device tensor, or
        [npts, 3] with the trailing columns ignored) for derivative
        ``orders`` (list of D-tuples; default value only) into ``out``
        ([nout, npts] device tensor, allocated if None; read it after
        ``wait``).  Returns the request id."""
        import torch
        if not orders:
            raise ValueError("orders must be given (one tuple of D derivative orders per output)")
        npts, cols = int(coords.shape[0]), int(coords.shape[1])
        coords = coords.contiguous()
        dim = len(orders[0])
        ords = np.ascontiguousarray(np.asarray(orders, dtype=np.int32).reshape(-1))
        nout = len(orders)
        f32 = out_dtype is not None and str(out_dtype) == "torch.float32"
        if out is None:
            out = torch.empty((nout, npts), dtype=torch.float32 if f32 else torch.float64, device=coords.device)
        flags = _lib.INTP_EVAL_OUT_F32 if str(out.dtype) == "torch.float32" else 0
        assert dim <= cols
        rid = C.c_int64()
        check(_lib.load().intp_serve_eval(self.h, int(spline), npts, C.cast(coords.data_ptr(), _dp), 1, cols, nout,
                                          ords.ctypes.data_as(C.POINTER(C.c_int32)), C.cast(out.data_ptr(), _dp),
                                          npts, 1, flags, C.byref(rid)))
        self._hold[rid.value] = (coords, out)
        return rid.value

    def poll(self, request: int):
        """(state, status): state is one of INTP_REQ_QUEUED/RUNNING/DONE/FAILED."""
        st, rc = C.c_int32(), C.c_int32()
        check(_lib.load().intp_serve_poll(self.h, int(request), C.byref(st), C.byref(rc)))
        return st.value, rc.value

    def wait(self, request: int):
        """Blocks until the request finishes; raises if it failed.  Returns
        the output tensor of an evaluation (None for a build)."""
        held = self._hold.pop(int(request), None)
        check(_lib.load().intp_serve_wait(self.h, int(request)))
        return held[1] if isinstance(held, tuple) else None

    def drain(self):
        """Blocks until every submitted request has finished."""
        check(_lib.load().intp_serve_drain(self.h))
        keep = {k: v for k, v in self._hold.items() if not isinstance(v, tuple)}
        self._hold = keep  # build inputs stay referenced by their ids until released

    def release(self, spline: int):
        check(_lib.load().intp_serve_release(self.h, int(spline)))
        self._hold.pop(int(spline), None)

    def spline(self, spline: int) -> InterpolationFunction:
        """The built spline as an InterpolationFunction (a new reference)."""
        h = C.c_void_p()
        check(_lib.load().intp_serve_spline(self.h, int(spline), C.byref(h)))
        return InterpolationFunction._wrap(h, self.ctx)

    def stats(self, reset: bool = False) -> dict:
        c = _lib.ServeCounters()
        check(_lib.load().intp_serve_stats(self.h, 1 if reset else 0, C.byref(c)))
        out = {}
        for k, name in ((_lib.INTP_REQ_BUILD, "build"), (_lib.INTP_REQ_EVAL, "eval")):
            t = c.type[k]
            out[name] = dict(completed=t.completed, failed=t.failed, bytes=t.bytes, p50_ms=t.latency_p50_ms,
                             p99_ms=t.latency_p99_ms, mean_ms=t.latency_mean_ms, max_ms=t.latency_max_ms)
        out.update(resident_splines=c.resident_splines, resident_bytes=c.resident_bytes, evictions=c.evictions,
                   admission_waits=c.admission_waits, first_submit_s=c.first_submit_s, last_done_s=c.last_done_s)
        return out
