"""The oracle against the compiled reference (oracle/_ref/libintp_ref.so,
the reference headers instantiated by oracle/ref_driver.cpp), on the same
inputs.  Skipped where the reference build is absent (the GPU host).

The oracle keeps the reference's arithmetic order and both are compiled
without FMA contraction, so the results are expected to be bit-identical;
the assertion allows 1e-15 relative to stay robust to libm differences.
"""
import itertools

import numpy as np
import pytest

from cases import interp_cases, rel_err

ref_mod = pytest.importorskip("oracle.oracle")
pytestmark = pytest.mark.skipif(not ref_mod.ref_available(), reason="reference build (oracle/_ref) absent")


def _configs():
    out = []
    rng = np.random.default_rng(2024)
    for D in (1, 2, 3):
        for k in range(0, 8):
            for per in itertools.product([False, True], repeat=D):
                if D == 3 and k not in (2, 3, 5, 7):
                    continue
                n = [int(rng.integers(max(k + 2, 6), 14)) for _ in range(D)]
                out.append((D, k, list(per), n, False))
    for D, k in ((1, 1), (1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7), (2, 3), (2, 4), (2, 7)):
        for per in itertools.product([False, True], repeat=D):
            n = [int(rng.integers(k + 3, 12)) for _ in range(D)]
            out.append((D, k, list(per), n, True))
    # 4D (uniform grids): the oracle's loops over any dimension count
    for k in range(0, 6):
        for per in ([False] * 4, [True, False, True, False], [True] * 4):
            n = [int(rng.integers(max(k + 2, 5), 8)) for _ in range(4)]
            out.append((4, k, per, n, False))
    return out


CONFIGS = _configs()


def _make(cfg, seed):
    D, k, per, n, nonuni = cfg
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(n)
    ranges, coords = [], []
    for d in range(D):
        lo = float(rng.uniform(-2, 0))
        hi = lo + float(rng.uniform(0.5, 3))
        ranges.append((lo, hi))
        if nonuni:
            m = n[d] + per[d]
            x = np.sort(rng.uniform(lo, hi, m))
            x[0], x[-1] = lo, hi
            coords.append(x)
        else:
            coords.append(None)
    return f, ranges, (coords if nonuni else None)


@pytest.mark.parametrize("cfg", CONFIGS, ids=["D%d_k%d_p%s_%s%s" % (c[0], c[1], "".join(str(int(p)) for p in c[2]),
                                                 "x".join(map(str, c[3])), "_nu" if c[4] else "")
                                              for c in CONFIGS])
def test_oracle_equals_reference(orc, cfg):
    D, k, per, n, nonuni = cfg
    f, ranges, coords = _make(cfg, hash((D, k, tuple(per), tuple(n))) & 0xFFFF)
    o = orc.OracleSpline(k, f, periodic=per, ranges=ranges, coords=coords)
    r = orc.RefSpline(k, f, periodic=per, ranges=ranges, coords=coords)
    rng = np.random.default_rng(77)
    lo = np.array([o.range(d)[0] for d in range(D)])
    hi = np.array([o.range(d)[1] for d in range(D)])
    for d in range(D):
        assert (o.range(d)) == (r.range(d))
    span = hi - lo
    q = lo + rng.uniform(-0.15, 1.15, (200, D)) * span
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)] + [[min(2, k + 1)] * D]
    got = o.eval(q, orders)
    want = r.eval(q, orders)
    for i in range(len(orders)):
        if np.all(want[i] == 0):
            assert np.all(got[i] == 0)
        else:
            assert rel_err(got[i], want[i]) < 1e-15, orders[i]


@pytest.mark.parametrize("case", interp_cases(), ids=[c["name"] for c in interp_cases()])
def test_reference_reproduces_mma(orc, case):
    """The compiled reference itself meets its own test bar on our
    transcription of the vectors (guards the transcription)."""
    r = orc.RefSpline(case["order"], case["f"], periodic=case["periodic"], ranges=case["ranges"],
                      coords=case["coords"])
    got = r.eval(case["points"], case["orders"], bounds=case["bounds"])[0]
    assert rel_err(got, case["expected"]) < case["tol"]


def test_reference_bounds_error(orc):
    c = [x for x in interp_cases() if x["name"] == "2d_cubic"][0]
    r = orc.RefSpline(3, c["f"], ranges=c["ranges"])
    with pytest.raises(ValueError):
        r.eval([[-1.0, 1.0]], bounds=True)


def test_oracle_bit_identical_3d_256_sample(orc):
    """A 3D order-3 build on a 48x40x36 grid (dim 0 periodic, the benchmark's
    boundary mix) with multithreaded solves on both sides: control-point
    values feed every query, so bit-identical outputs on 4096 points across
    the whole box pin the solve as well."""
    rng = np.random.default_rng(31)
    f = rng.standard_normal((48, 40, 36))
    per = [True, False, False]
    rg = [(0.0, 6.283185307179586), (0.0, 1.0), (0.0, 1.0)]
    o = orc.OracleSpline(3, f, periodic=per, ranges=rg, nthreads=8)
    r = orc.RefSpline(3, f, periodic=per, ranges=rg)
    q = np.stack([rng.uniform(0, 6.28, 4096), rng.uniform(0, 1, 4096), rng.uniform(0, 1, 4096)], 1)
    orders = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    a = o.eval(q, orders, nthreads=4)
    b = r.eval(q, orders, nthreads=4)
    assert np.array_equal(a, b)
