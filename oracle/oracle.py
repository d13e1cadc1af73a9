"""ctypes wrapper over oracle/liborc.so (the C restatement) and
oracle/_ref/libintp_ref.so (the reference headers instantiated by
ref_driver.cpp).  TEST INFRASTRUCTURE ONLY.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_PATH = os.path.join(_HERE, "liborc.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libintp_ref.so")
MAXD = 8

_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)


def build_oracle(ref: bool = True) -> None:
    """Runs oracle/Makefile (the restatement; the reference shim too when the
    reference checkout exists and ``ref`` is set)."""
    target = "all" if ref else "orc"
    subprocess.run(["make", "-s", "-C", _HERE, target], check=True)


def _lib_orc():
    if not os.path.exists(_ORC_PATH):
        build_oracle(ref=False)
    lib = C.CDLL(_ORC_PATH)
    lib.orc_new.restype = C.c_void_p
    lib.orc_delete.argtypes = [C.c_void_p]
    lib.orc_template_init.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), _i64p,
                                      _dp, _dp, C.POINTER(_dp)]
    lib.orc_template_init.restype = C.c_int
    lib.orc_bspline_init.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), _i64p,
                                     C.POINTER(_dp), _i64p, _dp]
    lib.orc_solve.argtypes = [C.c_void_p, _dp, C.c_int]
    lib.orc_solve.restype = C.c_int
    lib.orc_eval.argtypes = [C.c_void_p, _dp]
    lib.orc_eval.restype = C.c_double
    lib.orc_derivative.argtypes = [C.c_void_p, _dp, _i64p]
    lib.orc_derivative.restype = C.c_double
    lib.orc_eval_batch.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int, _i64p, _dp, C.c_int, C.c_int]
    lib.orc_eval_batch.restype = C.c_int
    lib.orc_nt.argtypes = [C.c_void_p, C.c_int]
    lib.orc_nt.restype = C.c_int64
    lib.orc_knots_ptr.argtypes = [C.c_void_p, C.c_int]
    lib.orc_knots_ptr.restype = _dp
    lib.orc_ctrl_ptr.argtypes = [C.c_void_p]
    lib.orc_ctrl_ptr.restype = _dp
    lib.orc_range.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.orc_base_spline_value.argtypes = [_dp, C.c_int64, C.c_double, C.c_int, C.c_int, _dp]
    lib.orc_band_alloc.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int]
    lib.orc_band_at.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
    lib.orc_band_at.restype = _dp
    lib.orc_band_lu.argtypes = [C.c_void_p]
    lib.orc_band_solve.argtypes = [C.c_void_p, _dp, C.c_int64]
    lib.orc_band_free.argtypes = [C.c_void_p]
    lib.orc_knots_uniform.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, _dp, _dp]
    lib.orc_knots_nonuniform.argtypes = [C.c_int, C.c_int, C.c_int64, _dp, _dp]
    lib.orc_knots_nonuniform.restype = C.c_int
    lib.orc_knots_count.argtypes = [C.c_int, C.c_int, C.c_int64]
    lib.orc_knots_count.restype = C.c_int64
    return lib


_ORC = None


def _orc():
    global _ORC
    if _ORC is None:
        _ORC = _lib_orc()
    return _ORC


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


_REF = None


def _ref():
    global _REF
    if _REF is None:
        lib = C.CDLL(_REF_PATH)
        lib.ref_build.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), _i64p, _dp, _dp,
                                  C.POINTER(_dp), _dp]
        lib.ref_build.restype = C.c_void_p
        lib.ref_eval.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int64, _i64p, _dp, C.c_int, C.c_int]
        lib.ref_eval.restype = C.c_int
        lib.ref_range.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        lib.ref_free.argtypes = [C.c_void_p]
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_bspline3_eval.argtypes = [C.c_int, C.POINTER(C.c_int), _i64p, C.POINTER(_dp), _i64p,
                                          _dp, C.c_int64, _dp, _i64p, _dp]
        lib.ref_bspline3_eval.restype = C.c_int
        _REF = lib
    return _REF


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_i64p)


class _Common:
    """Shared argument handling for OracleSpline and RefSpline."""

    def _prep(self, order, periodic, n, ranges, coords):
        self.dim = D = len(n)
        self.order = int(order)
        self.periodic = [bool(p) for p in (periodic or [False] * D)]
        self.n = np.ascontiguousarray(n, dtype=np.int64)
        lo = np.zeros(D)
        hi = np.zeros(D)
        coords = list(coords) if coords is not None else [None] * D
        self._coord_arrays = []
        cptrs = (_dp * MAXD)()
        for d in range(D):
            if coords[d] is not None:
                c = np.ascontiguousarray(coords[d], dtype=np.float64)
                if c.size != self.n[d] + self.periodic[d]:
                    raise ValueError("non-uniform coordinate count mismatch at dimension %d" % d)
                self._coord_arrays.append(c)
                cptrs[d] = _dptr(c)
            else:
                lo[d], hi[d] = ranges[d]
        self._lo, self._hi = lo, hi
        self._per = (C.c_int * MAXD)(*[int(p) for p in self.periodic])
        return cptrs


class OracleSpline(_Common):
    """InterpolationFunction<double, D, order> restated in C (build + eval)."""

    def __init__(self, order: int, f: np.ndarray, periodic: Optional[Sequence[bool]] = None,
                 ranges=None, coords=None, nthreads: int = 0):
        lib = _orc()
        f = np.ascontiguousarray(f, dtype=np.float64)
        cptrs = self._prep(order, periodic, f.shape, ranges, coords)
        self._h = lib.orc_new()
        rc = lib.orc_template_init(self._h, self.dim, self.order, self._per, _i64ptr(self.n),
                                   _dptr(self._lo), _dptr(self._hi), cptrs)
        if rc != 0:
            lib.orc_delete(self._h)
            self._h = None
            raise ValueError("oracle: invalid interpolation description")
        if lib.orc_solve(self._h, _dptr(f), int(nthreads)) != 0:
            raise MemoryError("oracle solve failed")

    @classmethod
    def from_knots(cls, order: int, knots: Sequence[np.ndarray], ctrl: np.ndarray,
                   periodic: Optional[Sequence[bool]] = None):
        """BSpline<double, D, order>(periodicity, ctrl, knot ranges...)."""
        self = cls.__new__(cls)
        lib = _orc()
        ctrl = np.ascontiguousarray(ctrl, dtype=np.float64)
        self.dim = D = ctrl.ndim
        self.order = order
        self.periodic = [bool(p) for p in (periodic or [False] * D)]
        self._per = (C.c_int * MAXD)(*[int(p) for p in self.periodic])
        self.n = np.ascontiguousarray(ctrl.shape, dtype=np.int64)
        ks = [np.ascontiguousarray(k, dtype=np.float64) for k in knots]
        kp = (_dp * MAXD)(*[_dptr(k) for k in ks])
        nk = np.array([k.size for k in ks], dtype=np.int64)
        self._h = lib.orc_new()
        lib.orc_bspline_init(self._h, D, order, self._per, _i64ptr(self.n), kp, _i64ptr(nk), _dptr(ctrl))
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            _orc().orc_delete(self._h)
            self._h = None

    def knots(self, d: int) -> np.ndarray:
        lib = _orc()
        nt = lib.orc_nt(self._h, d)
        return np.ctypeslib.as_array(lib.orc_knots_ptr(self._h, d), shape=(nt,)).copy()

    def control_points(self) -> np.ndarray:
        p = _orc().orc_ctrl_ptr(self._h)
        return np.ctypeslib.as_array(p, shape=(int(np.prod(self.n)),)).reshape(tuple(self.n)).copy()

    def range(self, d: int):
        lo, hi = C.c_double(), C.c_double()
        _orc().orc_range(self._h, d, C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def eval(self, pts: np.ndarray, orders=None, nthreads: int = 1, bounds: bool = False) -> np.ndarray:
        """pts [npts, D]; orders [nderiv, D] (default: value only).
        Returns [nderiv, npts]."""
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, self.dim))
        orders = np.zeros((1, self.dim), np.int64) if orders is None else \
            np.ascontiguousarray(np.asarray(orders, dtype=np.int64).reshape(-1, self.dim))
        out = np.empty((orders.shape[0], pts.shape[0]))
        rc = _orc().orc_eval_batch(self._h, pts.shape[0], _dptr(pts), orders.shape[0], _i64ptr(orders),
                                   _dptr(out), int(nthreads), int(bounds))
        if rc != 0:
            raise ValueError("Given coordinate out of interpolation function range!")
        return out


class RefSpline(_Common):
    """The compiled reference: intp::InterpolationFunction<double, D, order>."""

    def __init__(self, order: int, f: np.ndarray, periodic: Optional[Sequence[bool]] = None,
                 ranges=None, coords=None):
        lib = _ref()
        f = np.ascontiguousarray(f, dtype=np.float64)
        cptrs = self._prep(order, periodic, f.shape, ranges, coords)
        any_nu = any(c is not None for c in (coords or []))
        self._h = lib.ref_build(self.dim, self.order, self._per, _i64ptr(self.n), _dptr(self._lo),
                                _dptr(self._hi), cptrs if any_nu else None, _dptr(f))
        if not self._h:
            raise ValueError("reference build failed: %s" % lib.ref_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            _ref().ref_free(self._h)
            self._h = None

    def range(self, d: int):
        lo, hi = C.c_double(), C.c_double()
        _ref().ref_range(self._h, d, C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def eval(self, pts: np.ndarray, orders=None, nthreads: int = 1, bounds: bool = False) -> np.ndarray:
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, self.dim))
        orders = np.zeros((1, self.dim), np.int64) if orders is None else \
            np.ascontiguousarray(np.asarray(orders, dtype=np.int64).reshape(-1, self.dim))
        out = np.empty((orders.shape[0], pts.shape[0]))
        rc = _ref().ref_eval(self._h, pts.shape[0], _dptr(pts), orders.shape[0], _i64ptr(orders),
                             _dptr(out), int(nthreads), int(bounds))
        if rc != 0:
            raise ValueError("Given coordinate out of interpolation function range!")
        return out


def ref_bspline3(knots, ctrl, pts, order_r, periodic=None) -> np.ndarray:
    """Reference BSpline<double, D, 3>::derivative_at on explicit knots."""
    ctrl = np.ascontiguousarray(ctrl, dtype=np.float64)
    D = ctrl.ndim
    per = (C.c_int * MAXD)(*[int(p) for p in (periodic or [False] * D)])
    n = np.ascontiguousarray(ctrl.shape, dtype=np.int64)
    ks = [np.ascontiguousarray(k, dtype=np.float64) for k in knots]
    kp = (_dp * MAXD)(*[_dptr(k) for k in ks])
    nk = np.array([k.size for k in ks], dtype=np.int64)
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, D))
    r = np.ascontiguousarray(order_r, dtype=np.int64)
    out = np.empty(pts.shape[0])
    rc = _ref().ref_bspline3_eval(D, per, _i64ptr(n), kp, _i64ptr(nk), _dptr(ctrl), pts.shape[0],
                                  _dptr(pts), _i64ptr(r), _dptr(out))
    if rc != 0:
        raise ValueError(_ref().ref_last_error().decode())
    return out


class Oracle:
    """Namespace of low-level oracle entry points for unit tests."""

    @staticmethod
    def knots_uniform(order, periodic, n, xmin, xmax):
        lib = _orc()
        nt = lib.orc_knots_count(order, int(periodic), n)
        t = np.zeros(nt)
        dx = C.c_double()
        lib.orc_knots_uniform(order, int(periodic), n, xmin, xmax, _dptr(t), C.byref(dx))
        return t, dx.value

    @staticmethod
    def knots_nonuniform(order, periodic, n, xs):
        lib = _orc()
        nt = lib.orc_knots_count(order, int(periodic), n)
        t = np.zeros(nt)
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        if lib.orc_knots_nonuniform(order, int(periodic), n, _dptr(xs), _dptr(t)) != 0:
            raise ValueError("invalid non-uniform coordinates")
        return t

    @staticmethod
    def basis(knots, mu, x, order, spline_order=None):
        lib = _orc()
        t = np.ascontiguousarray(knots, dtype=np.float64)
        out = np.zeros(order + 1)
        lib.orc_base_spline_value(_dptr(t), mu, x, order, order if spline_order is None else spline_order,
                                  _dptr(out))
        return out

    @staticmethod
    def band_solve(dense: np.ndarray, p: int, q: int, extended: bool, b: np.ndarray) -> np.ndarray:
        """LU-factor the band (+ cyclic strips) of ``dense`` and solve for b."""
        lib = _orc()
        n = dense.shape[0]
        buf = C.create_string_buffer(256)
        lib.orc_band_alloc(buf, n, p, q, int(extended))
        for i in range(n):
            for j in range(n):
                v = dense[i, j]
                if v == 0.0:
                    continue
                in_main = (j <= i + q) and (i <= j + p)
                in_side = extended and ((j >= max(n - p, i + q + 1)) or (i >= max(n - q, j + p + 1)))
                if not (in_main or in_side):
                    raise ValueError("entry (%d,%d) outside the band storage" % (i, j))
                lib.orc_band_at(buf, i, j)[0] = v
        lib.orc_band_lu(buf)
        x = np.ascontiguousarray(b, dtype=np.float64).copy()
        lib.orc_band_solve(buf, _dptr(x), 1)
        lib.orc_band_free(buf)
        return x
