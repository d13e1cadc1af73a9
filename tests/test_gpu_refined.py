"""Refined-FP32 form (intp_build_refined): FP32 control points solved with
the FP32-rounded LU factors plus FP64-residual refinement sweeps, evaluated
with FP32 weights and tables.  Checked against the FP64 oracle with these
error measures (at most 1e-6):

    max_rel = max|got - want| / max|want|,   rms_rel = rms(got - want) / rms(want)

Bars once at least one refinement sweep has run, for data shaped like the
benchmark workloads (a smooth function plus 1e-3 noise): control points
within 1e-6 (max) of the FP64 solve; values within 1e-6 norm-wise (rms_rel,
the norm-wise relative error the FP64 path is held to at 1e-12) and 5e-6
max; first partials, evaluated from the gradient table's FP32 differences
of the corrected control points, within 1e-6 max.  White-noise data (a
stress case: its control points grow and cancel, so FP32 storage of them
costs more) is held to 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BAR = 1e-6


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def errs(got, want):
    d = np.asarray(got, np.float64) - want
    return float(np.max(np.abs(d)) / np.max(np.abs(want))), float(np.sqrt(np.mean(d * d) / np.mean(want * want)))


def _points(rng, o, D, n):
    lo = np.array([o.range(d)[0] for d in range(D)])
    hi = np.array([o.range(d)[1] for d in range(D)])
    return lo + rng.uniform(0, 1, (n, D)) * (hi - lo)


CASES = [
    # shape, order, periodic, non-uniform dims
    ((40, 36, 32), 5, [False, False, False], []),
    ((48, 40, 36), 3, [True, False, False], []),
    ((100, 90), 4, [True, True], []),
    ((60, 50), 3, [False, True], [0]),
    ((8192,), 6, [True], []),        # periodic order >= 5: FP64 factors for that axis
    ((20000,), 3, [False], []),      # longer than an FP32 LDS tile: in-place HBM solve
    ((30, 28, 26), 6, [True, False, True], []),
]


def _smooth(rng, shape, per):
    """sum of products of low sinusoids (periodic in the periodic dims) + 1e-3 noise"""
    D = len(shape)
    axes = [np.arange(n) / n if p else np.linspace(0, 1, n) for n, p in zip(shape, per)]
    grid = np.meshgrid(*axes, indexing="ij")
    f = np.zeros(shape)
    for m in range(3):
        term = rng.uniform(0.5, 1.5)
        for d in range(D):
            term = term * np.cos(2 * np.pi * (m + 1) * grid[d] + rng.uniform(0, 6.3))
        f += term
    return f + 1e-3 * rng.standard_normal(shape)


@pytest.mark.parametrize("data", ["smooth", "noise"])
@pytest.mark.parametrize("shape,k,per,nonuni", CASES, ids=lambda v: str(v))
def test_gpu_refined_against_fp64_oracle(amd, orc, shape, k, per, nonuni, data):
    rng = np.random.default_rng(sum(shape) * 7 + k)
    D = len(shape)
    f = _smooth(rng, shape, per) if data == "smooth" else rng.standard_normal(shape)
    value_bar = BAR if data == "smooth" else 10 * BAR
    rg = [(0.0, 1.0 + 0.5 * d) for d in range(D)]
    coords = [None] * D
    for d in nonuni:
        c = np.cumsum(rng.uniform(0.5, 1.5, shape[d] + (1 if per[d] else 0)))
        coords[d] = c - c[0]
    o = orc.OracleSpline(k, f, periodic=per, ranges=rg, coords=coords)
    want_ctrl = o.control_points()
    q = _points(rng, o, D, 4000)
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)]
    want = o.eval(q, orders, nthreads=8)
    seen = []
    for sweeps in (0, 1, 2, 3):
        g = amd.InterpolationFunction(k, f, periodic=per, ranges=rg, coords=coords, precision="f32_refined",
                                      sweeps=sweeps)
        assert g.precision == "f32_refined" and g.refine_sweeps == sweeps
        c_max, c_rms = errs(g.control_points(), want_ctrl)
        got = g.eval(q, orders)
        v_max, v_rms = errs(got[0], want[0])
        seen.append((sweeps, c_max, v_max))
        print("sweeps %d: ctrl max %.2e rms %.2e | value max %.2e rms %.2e" % (sweeps, c_max, c_rms, v_max, v_rms))
        # the host scalar path: the FP32 control points for values, the
        # gradient table's differences for first partials
        host = g.eval_host(q, orders)
        for i in range(1, D + 1):
            d_max, d_rms = errs(got[i], want[i])
            h_max, _ = errs(host[i], want[i])
            print("  d/dx%d: max %.2e rms %.2e | host max %.2e" % (i - 1, d_max, d_rms, h_max))
            if sweeps >= 1:
                assert d_max <= value_bar and h_max <= value_bar, (sweeps, i, d_max, d_rms, h_max)
        if sweeps >= 1:
            assert errs(host[0], want[0])[1] <= value_bar
        if sweeps >= 1:
            assert c_max <= BAR and v_rms <= value_bar and v_max <= 5 * value_bar, seen
    # refinement never makes the control points worse than the plain FP32 solve
    assert seen[-1][1] <= max(seen[0][1], 2e-7)


def test_gpu_refined_f32_outputs_and_grouping(amd):
    import torch
    rng = np.random.default_rng(41)
    f = rng.standard_normal((64, 48, 40))
    rg = [(0, 2), (0, 1), (-1, 1)]
    g = amd.InterpolationFunction(5, f, periodic=[False, True, False], ranges=rg, precision="f32_refined",
                                  sweeps=2)
    q = np.stack([rng.uniform(r[0], r[1], (1 << 20) + 9) for r in rg])
    qd = torch.from_numpy(q).cuda()
    orders = [[0, 0, 0], [0, 1, 0], [2, 0, 1]]
    out64 = g.eval(qd, orders, group=True)             # grouped, FP64 results of FP32 arithmetic
    out32 = g.eval(qd, orders, out_dtype=torch.float32, group=True)
    assert np.array_equal(out32.cpu().numpy(), out64.cpu().numpy().astype(np.float32))
    # caller-order evaluation gives the same bits
    import ctypes as C
    from bsplineinterpolation_amd import _lib
    plain = torch.empty_like(out64)
    op = np.ascontiguousarray(np.asarray(orders, np.int32))
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().intp_eval(g.h, q.shape[1], C.cast(qd.data_ptr(), dp), q.shape[1], 1, 3,
                                     op.ctypes.data_as(C.POINTER(C.c_int32)), C.cast(plain.data_ptr(), dp),
                                     q.shape[1], 1, _lib.INTP_EVAL_UNGROUPED,
                                     C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    assert torch.equal(plain, out64)


def test_gpu_refined_template_and_validation(amd, orc):
    rng = np.random.default_rng(5)
    t = amd.InterpolationFunctionTemplate(4, (50, 44), periodic=[True, False], ranges=[(0, 1), (0, 1)])
    f = rng.standard_normal((50, 44))
    a = t.interpolate(f, precision="f32_refined", sweeps=1)
    b = amd.InterpolationFunction(4, f, periodic=[True, False], ranges=[(0, 1), (0, 1)], precision="f32_refined",
                                  sweeps=1)
    assert np.array_equal(a.control_points(), b.control_points())
    o = orc.OracleSpline(4, f, periodic=[True, False], ranges=[(0, 1), (0, 1)])
    assert errs(a.control_points(), o.control_points())[0] <= BAR
    with pytest.raises(amd.IntpError):
        t.interpolate(f, precision="f32_refined", sweeps=4)
    with pytest.raises(amd.IntpError):
        t.interpolate(f, precision="f16")
