"""Portable, seeded synthetic inputs for the benchmark workloads and tests.

Counter-based SplitMix64: value ``i`` of stream ``s`` under ``seed`` is
``mix(key(seed, s) + (i + 1) * 0x9E3779B97F4A7C15)``; uniforms take the top
53 bits, Gaussians use Box-Muller on consecutive uniform pairs.  Any point
of any stream can be regenerated independently, in any language, which the
fixtures and the CPU baseline rely on.  ``std::*_distribution`` is never
used: its output is implementation-defined.
"""
from __future__ import annotations

import math

import numpy as np

SEED = 0x20261202
STREAM_NOISE = 0
STREAM_QUERY = 1
STREAM_SPEC = 2

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_SKEY = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """``count`` doubles in [0, 1) from positions offset .. offset+count-1."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream) * _SKEY)],
                            dtype=np.uint64))[0]
        out = np.empty(count, dtype=np.float64)
        chunk = 1 << 22
        for b in range(0, count, chunk):
            e = min(count, b + chunk)
            i = np.arange(offset + b + 1, offset + e + 1, dtype=np.uint64)
            z = _mix(key + i * _GOLD)
            out[b:e] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out


def normal(seed: int, stream: int, count: int, sigma: float = 1.0) -> np.ndarray:
    """Box-Muller: pair j uses uniforms 2j, 2j+1 of the stream."""
    m = (count + 1) // 2
    u = uniform(seed, stream, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return sigma * z[:count]


def grid_axis(lo: float, hi: float, n: int, periodic: bool) -> np.ndarray:
    """Sample sites: n points on [lo, hi] (clamped) or n points of [lo, hi)
    with spacing (hi - lo)/n (periodic, INTP_PERIODIC_NO_DUMMY_POINT)."""
    if periodic:
        return lo + (hi - lo) * np.arange(n) / n
    return lo + (hi - lo) * np.arange(n) / (n - 1)


def queries(seed: int, lo, hi, count: int, stream: int = STREAM_QUERY) -> np.ndarray:
    """``count`` points, uniform in the box [lo, hi); SoA array [D, count].
    Coordinate d of point i uses position d*count + i of the stream."""
    D = len(lo)
    u = uniform(seed, stream, D * count).reshape(D, count)
    for d in range(D):
        u[d] = lo[d] + (hi[d] - lo[d]) * u[d]
    return u


# ---------------------------------------------------------------------------
# the named workloads
# ---------------------------------------------------------------------------

def workload_2d_64_cubic(seed: int = SEED, nq: int = 1_000_000):
    """2D order-3 clamped FP64 on a 64x64 mesh over [0,1]^2,
    f = sin(2 pi x) cos(3 pi y) + N(0, 1e-2^2); nq uniform queries."""
    x = grid_axis(0.0, 1.0, 64, False)
    y = grid_axis(0.0, 1.0, 64, False)
    f = np.sin(2 * math.pi * x)[:, None] * np.cos(3 * math.pi * y)[None, :]
    f = f + normal(seed, STREAM_NOISE, f.size, 1e-2).reshape(f.shape)
    q = queries(seed, [0.0, 0.0], [1.0, 1.0], nq)
    return dict(order=3, f=f, periodic=[False, False], ranges=[(0.0, 1.0), (0.0, 1.0)], queries=q,
                orders=[[0, 0], [1, 0]])


def workload_3d_256_cubic(seed: int = SEED, n: int = 256, nq: int = 1 << 26, nq_gen: int = None):
    """3D order-3 FP64 on n^3 over [0,2pi) x [0,1] x [0,1], dim 0 periodic,
    f = sin x exp(-(y-1/2)^2/0.05) cos(4 pi z) + N(0, 1e-3^2); nq uniform
    queries for value and the three first partials.  ``nq_gen`` (<= nq)
    generates only the leading queries (same values as the full set)."""
    x = grid_axis(0.0, 2 * math.pi, n, True)
    y = grid_axis(0.0, 1.0, n, False)
    z = grid_axis(0.0, 1.0, n, False)
    f = (np.sin(x)[:, None, None] * np.exp(-((y - 0.5) ** 2) / 0.05)[None, :, None]
         * np.cos(4 * math.pi * z)[None, None, :])
    f = f + normal(seed, STREAM_NOISE, f.size, 1e-3).reshape(f.shape)
    lo, hi = [0.0, 0.0, 0.0], [2 * math.pi, 1.0, 1.0]
    m = nq if nq_gen is None else nq_gen
    u = uniform(seed, STREAM_QUERY, 3 * nq if nq_gen is None else 0)
    if nq_gen is None:
        q = u.reshape(3, nq)
        for d in range(3):
            q[d] = lo[d] + (hi[d] - lo[d]) * q[d]
    else:
        q = np.empty((3, m))
        for d in range(3):
            q[d] = lo[d] + (hi[d] - lo[d]) * uniform(seed, STREAM_QUERY, m, offset=d * nq)
    return dict(order=3, f=f, periodic=[True, False, False], ranges=[(0.0, 2 * math.pi), (0.0, 1.0), (0.0, 1.0)],
                queries=q, orders=[[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def workload_3d_256_quintic(seed: int = SEED, n: int = 256, nq: int = 1 << 26):
    """3D order-5 on n^3 over [0,2pi] x [0,1] x [0,1], all dimensions
    clamped (x_i = 2 pi i / (n-1)), the same value formula and noise as
    workload_3d_256_cubic; nq uniform queries, value only.  The
    refined-FP32 form is measured on it."""
    x = grid_axis(0.0, 2 * math.pi, n, False)
    y = grid_axis(0.0, 1.0, n, False)
    z = grid_axis(0.0, 1.0, n, False)
    f = (np.sin(x)[:, None, None] * np.exp(-((y - 0.5) ** 2) / 0.05)[None, :, None]
         * np.cos(4 * math.pi * z)[None, None, :])
    f = f + normal(seed, STREAM_NOISE, f.size, 1e-3).reshape(f.shape)
    lo, hi = [0.0, 0.0, 0.0], [2 * math.pi, 1.0, 1.0]
    q = uniform(seed, STREAM_QUERY, 3 * nq).reshape(3, nq)
    for d in range(3):
        q[d] = lo[d] + (hi[d] - lo[d]) * q[d]
    return dict(order=5, f=f, periodic=[False, False, False], ranges=[(0.0, 2 * math.pi), (0.0, 1.0), (0.0, 1.0)],
                queries=q, orders=[[0, 0, 0]])


def workload_1d_periodic_sextic(seed: int = SEED, n: int = 8192, nq: int = 1 << 20):
    """1D periodic order 6, n values on [0, 2pi): sin x + 0.3 cos 5x +
    N(0, 1e-3^2), nq uniform queries (value).  Its periodic order-6 solve
    keeps FP64 factors in the refined form."""
    x = grid_axis(0.0, 2 * math.pi, n, True)
    f = np.sin(x) + 0.3 * np.cos(5 * x) + normal(seed, STREAM_NOISE, n, 1e-3)
    q = 2 * math.pi * uniform(seed, STREAM_QUERY, nq).reshape(1, nq)
    return dict(order=6, f=f, periodic=[True], ranges=[(0.0, 2 * math.pi)], queries=q, orders=[[0]])


def workload_batch_mixed(seed: int = SEED, count: int = 4096, queries_per_spline: int = 256):
    """The batched workload: ``count`` FP64 splines, half 1D with lengths
    U{16..4096}, 3/8 2D with sides U{16..256}, 1/8 3D with sides U{8..64};
    orders U{1..5}; each dimension periodic with probability 1/4; ranges
    [0, 1].  Values: a sum of 3 sinusoids with seeded amplitudes, frequencies
    and phases plus N(0, 1e-2^2); ``queries_per_spline`` uniform points each
    (value).  Returns specs, concatenated values, points [npts, 3] and
    point offsets."""
    n1, n2 = count // 2, (count * 3) // 8
    u = uniform(seed, STREAM_SPEC, 16 * count).reshape(count, 16)

    def pick(v, lo, hi):
        return lo + int(v * (hi - lo + 1)) if v < 1.0 else hi

    specs, vals, pts, offs = [], [], [], [0]
    for i in range(count):
        r = u[i]
        D = 1 if i < n1 else (2 if i < n1 + n2 else 3)
        lo_n, hi_n = ((16, 4096), (16, 256), (8, 64))[D - 1]
        shape = tuple(pick(r[d], lo_n, hi_n) for d in range(D))
        order = pick(r[3], 1, 5)
        periodic = [bool(r[4 + d] < 0.25) for d in range(D)]
        # a periodic dimension needs n >= 2*floor(k/2) + 2
        shape = tuple(max(s, 2 * (order // 2) + 2) if p else max(s, order + 1) for s, p in zip(shape, periodic))
        specs.append(dict(order=order, shape=shape, periodic=periodic, ranges=[(0.0, 1.0)] * D))
        axes = [np.arange(n) / n if p else np.linspace(0.0, 1.0, n) for n, p in zip(shape, periodic)]
        grid = np.meshgrid(*axes, indexing="ij")
        f = np.zeros(shape)
        for m in range(3):
            amp = 0.5 + r[7 + m]
            arg = np.zeros(shape)
            for d in range(D):
                arg = arg + (1.0 + 5.0 * r[10 + (m + d) % 3]) * grid[d]
            f = f + amp * np.sin(2 * math.pi * arg + 6.28 * r[13 + m])
        vals.append(f.ravel())
        offs.append(offs[-1] + queries_per_spline)
    pts = [uniform(seed, STREAM_QUERY, 3 * offs[-1]).reshape(offs[-1], 3)]
    values = np.concatenate(vals)
    values = values + normal(seed, STREAM_NOISE, values.size, 1e-2)
    return dict(specs=specs, values=values, points=np.concatenate(pts), point_offsets=np.asarray(offs, np.int64))


STREAM_SERVE = 3


def workload_serving(seed: int = SEED, rate: float = 1e3, seconds: float = 2.0, count: int = None,
                     phase: int = 0, recent: int = 16):
    """Requests of the serving workload: Poisson arrivals at ``rate`` per
    second for ``seconds`` (or exactly ``count`` requests, all arriving at
    t = 0 when rate is None).  30% builds: 2D with sides U{32..1024} or
    3D with sides U{32..192} (even odds), order U{3,4,5}, all dimensions
    clamped on [0, 1], FP64 or refined FP32 (even odds, 1 sweep).  70%
    evaluations of log-uniform 1e3..1e6 points, half of them with every
    first partial next to the value, against one of the ``recent`` latest
    builds.  Each request: dict(t, kind, ...); uniforms come from stream
    STREAM_SERVE at offset 8 * (request index + 2^24 * phase)."""
    n = count if count is not None else int(rate * seconds * 1.5) + 16
    u = uniform(seed, STREAM_SERVE, 8 * n, offset=8 * (phase << 24)).reshape(n, 8)
    reqs, builds, t = [], [], 0.0
    for i in range(n):
        r = u[i]
        if rate is not None:
            t += -math.log1p(-r[0]) / rate
            if count is None and t > seconds:
                break
        if r[1] < 0.3 or not builds:
            D = 2 if r[2] < 0.5 else 3
            lo, hi = (32, 1024) if D == 2 else (32, 192)
            shape = tuple(lo + min(int(r[3 + d] * (hi - lo + 1)), hi - lo) for d in range(D))
            order = 3 + min(int(r[6] * 3), 2)
            precision = "f64" if r[7] < 0.5 else "f32_refined"
            builds.append(len(reqs))
            reqs.append(dict(t=t if rate is not None else 0.0, kind="build", shape=shape, order=order,
                             precision=precision))
        else:
            npts = int(round(10 ** (3 + 3 * r[2])))
            target = builds[-1 - min(int(r[3] * recent), len(builds) - 1)]
            reqs.append(dict(t=t if rate is not None else 0.0, kind="eval", target=target, npts=npts,
                             derivs=bool(r[4] < 0.5)))
    return reqs
