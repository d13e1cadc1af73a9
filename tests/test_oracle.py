"""Pins the CPU oracle before anything is checked against it.

* the reference's own known-answer vectors (MMA, bspline-test, band solver);
* what mathematics fixes independently of the reference's code: scipy's
  B-spline routines on the same knots, brute-force span search, partition of
  unity, polynomial reproduction, periodic invariance, dense linear solves.
Each check is chosen so that a dropped term, a wrong index or sign, or a
transposed operand in the oracle fails it.
"""
import math

import numpy as np
import pytest
import scipy.interpolate as si

from cases import bspline_cases, interp_cases, rel_err

CASES = interp_cases()
BS_CASES = bspline_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_mma_interpolation(orc, case):
    s = orc.OracleSpline(case["order"], case["f"], periodic=case["periodic"], ranges=case["ranges"],
                         coords=case["coords"])
    got = s.eval(case["points"], case["orders"], bounds=case["bounds"])[0]
    assert rel_err(got, case["expected"]) < case["tol"], case["cite"]


@pytest.mark.parametrize("case", BS_CASES, ids=[c["name"] for c in BS_CASES])
def test_oracle_mma_bspline(orc, case):
    s = orc.OracleSpline.from_knots(3, case["knots"], case["ctrl"], periodic=case["periodic"])
    got = s.eval(case["points"], case["orders"])[0]
    assert rel_err(got, case["expected"]) < case["tol"], case["cite"]


def test_oracle_at_out_of_range_raises(orc):
    """interpolation-test.cpp:159-164: at(-1, 1) throws std::domain_error."""
    c = [x for x in CASES if x["name"] == "2d_cubic"][0]
    s = orc.OracleSpline(3, c["f"], ranges=c["ranges"])
    with pytest.raises(ValueError):
        s.eval([[-1.0, 1.0]], bounds=True)
    s.eval([[0.0, 4.0]], bounds=True)  # both ends inclusive


def _dense_cyclic(n, diag):
    a = np.zeros((n, n))
    for off, v in diag.items():
        for i in range(n):
            a[i, (i + off) % n] = v
    return a


@pytest.mark.parametrize("p,q,diag", [
    (1, 1, {-1: 1 / 8, 0: 3 / 4, 1: 1 / 8}),
    (2, 2, {-2: 1 / 384, -1: 19 / 96, 0: 115 / 192, 1: 19 / 96, 2: 1 / 384}),
    (1, 3, {-1: 2889 / 16000, 0: 1701 / 3200, 1: 33 / 128, 2: 729 / 20000, 3: 9 / 20000}),
])
def test_oracle_extended_band_solver(orc, p, q, diag):
    """band-matrix-and-solver-test.cpp:68-106 (seeded RHS instead of the
    time seed), checked against a dense numpy solve as well as the
    residual bound of the reference test."""
    n = 64
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, n)
    a = _dense_cyclic(n, diag)
    x = orc.Oracle.band_solve(a, p, q, True, b)
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-10
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-13, atol=1e-13)


def test_oracle_band_solver_laplacian(orc):
    """band-matrix-and-solver-test.cpp:54-66."""
    n = 64
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = -2
        if i > 0:
            a[i, i - 1] = 1
        if i < n - 1:
            a[i, i + 1] = 1
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x = orc.Oracle.band_solve(a, 1, 1, False, b)
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-10
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-12)


def test_oracle_band_solver_nonsymmetric(orc):
    """A non-symmetric, unequal-bandwidth band (p=1, q=3): catches a
    transposed factor or a swapped p/q in the substitution."""
    n = 40
    rng = np.random.default_rng(11)
    a = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - 1), min(n, i + 4)):
            a[i, j] = rng.uniform(0.1, 1.0)
        a[i, i] = 5.0
    b = rng.uniform(-1, 1, n)
    x = orc.Oracle.band_solve(a, 1, 3, False, b)
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------
# knots
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [1, 3, 5])
def test_oracle_uniform_knots_are_not_a_knot(orc, k):
    """Odd order, clamped: the knot vector is scipy's not-a-knot vector."""
    n = 17
    t, dx = orc.Oracle.knots_uniform(k, False, n, 0.0, 2.0)
    x = np.linspace(0.0, 2.0, n)
    t_ref = np.r_[[x[0]] * (k + 1), x[(k + 1) // 2: -(k + 1) // 2], [x[-1]] * (k + 1)]
    assert np.allclose(t, t_ref, rtol=0, atol=1e-15)
    assert dx == pytest.approx(2.0 / (n - 1), rel=1e-15)


@pytest.mark.parametrize("k", [0, 1, 2, 3, 4, 5, 6, 7])
def test_oracle_periodic_uniform_knots(orc, k):
    """Periodic: equally spaced with spacing P/n, data sites at knots (odd k)
    or at knot midpoints (even k); range covers exactly one period."""
    n, a, b = 12, 1.0, 4.0
    t, dx = orc.Oracle.knots_uniform(k, True, n, a, b)
    assert np.allclose(np.diff(t), (b - a) / n, rtol=1e-13)
    s = orc.OracleSpline(k, np.arange(n, dtype=float) if k else np.arange(n, dtype=float),
                         periodic=[True], ranges=[(a, b)])
    lo, hi = s.range(0)
    assert hi - lo == pytest.approx(b - a, rel=1e-14)
    assert lo == pytest.approx(a if k % 2 else a - dx / 2, abs=1e-14)


def test_oracle_nonuniform_knots_moving_average(orc):
    """Clamped non-uniform: interior knots are order-window means of the
    node coordinates (Interpolation.hpp:418-447)."""
    xs = np.array([0.0, 0.3, 1.1, 1.5, 2.8, 3.0, 4.4, 5.0])
    k = 3
    t = orc.Oracle.knots_nonuniform(k, False, xs.size, xs)
    inner = [xs[i - k: i].mean() for i in range(k + 1, xs.size)]
    assert np.allclose(t[k + 1: xs.size], inner, rtol=1e-15)
    assert np.all(t[: k + 1] == xs[0]) and np.all(t[xs.size:] == xs[-1])
    with pytest.raises(ValueError):
        orc.Oracle.knots_nonuniform(k, False, 4, np.array([0.0, 1.0, 1.0, 2.0]))


# ---------------------------------------------------------------------------
# basis, span
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [0, 1, 2, 3, 4, 5, 6, 7])
def test_oracle_basis_matches_scipy(orc, k):
    """Cox-de Boor against scipy's design matrix on a non-uniform clamped
    knot vector; also partition of unity."""
    rng = np.random.default_rng(k)
    inner = np.sort(rng.uniform(0.1, 0.9, 9))
    t = np.r_[[0.0] * (k + 1), inner, [1.0] * (k + 1)]
    nb = t.size - k - 1
    for x in rng.uniform(0.0, 1.0 - 1e-9, 25):
        mu = int(np.searchsorted(t, x, side="right") - 1)
        mu = min(max(mu, k), t.size - k - 2)
        b = orc.Oracle.basis(t, mu, x, k)
        dm = si.BSpline.design_matrix(np.array([x]), t, k).toarray()[0]
        assert np.allclose(b, dm[mu - k: mu + 1], rtol=1e-13, atol=1e-15)
        assert abs(b.sum() - 1.0) < 1e-14
        assert dm.shape[0] == nb


def test_oracle_reduced_order_basis_right_aligned(orc):
    """spline_order < order: the lower-order basis sits right-aligned
    (BSpline.hpp:71-76); compare with scipy at order k-1."""
    k = 4
    t = np.r_[[0.0] * 5, 0.2, 0.35, 0.6, 0.8, [1.0] * 5]
    x = 0.47
    mu = int(np.searchsorted(t, x, side="right") - 1)
    b = orc.Oracle.basis(t, mu, x, k, k - 1)
    dm = si.BSpline.design_matrix(np.array([x]), t[1:-1], k - 1).toarray()[0]
    assert b[0] == 0.0
    assert np.allclose(b[1:], dm[mu - 1 - (k - 1): mu], rtol=1e-13, atol=1e-15)


def test_oracle_span_search_brute_force(orc):
    """Span search (hint or binary search) equals a linear scan, including
    exact knot hits, ends and out-of-range extrapolation."""
    rng = np.random.default_rng(5)
    for k in (1, 2, 3, 5):
        for periodic in (False, True):
            n = 11
            f = rng.standard_normal(n)
            s = orc.OracleSpline(k, f, periodic=[periodic], ranges=[(-1.0, 2.0)])
            t = s.knots(0)
            lo, hi = s.range(0)
            pts = np.r_[rng.uniform(lo - 0.7, hi + 0.7, 60), t[k: t.size - k], lo, hi]
            # reconstruct the span from a value evaluation: compare with a
            # direct sum over the brute-force span
            c = s.control_points().ravel()
            for x in pts:
                xx = x
                if periodic:
                    P = hi - lo
                    xx = lo + math.fmod(x - lo, P) + (P if x < lo else 0.0)
                mu = k
                for m in range(k, t.size - k - 1):
                    if t[m] <= xx:
                        mu = m
                b = orc.Oracle.basis(t, mu, xx, k)
                idx = [(mu - k + j) % n if periodic else mu - k + j for j in range(k + 1)]
                want = float(np.dot(b, c[idx]))
                got = s.eval([[x]])[0, 0]
                assert got == pytest.approx(want, rel=1e-13, abs=1e-13)


# ---------------------------------------------------------------------------
# end-to-end interpolation against scipy / linear algebra
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7])
def test_oracle_clamped_1d_matches_make_interp_spline(orc, k):
    """Same knots, same data sites -> the interpolant is unique, so scipy's
    make_interp_spline(t=...) must give the same coefficients."""
    n = 23
    rng = np.random.default_rng(100 + k)
    f = rng.standard_normal(n)
    s = orc.OracleSpline(k, f, ranges=[(0.0, 3.0)])
    t = s.knots(0)
    x = np.linspace(0.0, 3.0, n)
    sp = si.make_interp_spline(x, f, k=k, t=t)
    assert np.allclose(s.control_points(), sp.c, rtol=1e-12, atol=1e-12)
    q = rng.uniform(0.0, 3.0, 50)
    assert np.allclose(s.eval(q[:, None])[0], sp(q), rtol=1e-12, atol=1e-12)
    for r in range(1, k + 1):
        assert np.allclose(s.eval(q[:, None], [[r]])[0], sp(q, nu=r), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("k", [1, 3, 5, 7])
def test_oracle_periodic_odd_matches_scipy(orc, k):
    """Odd order, periodic, uniform: scipy's periodic knots coincide with the
    reference's, so values must agree."""
    n = 20
    rng = np.random.default_rng(200 + k)
    f = rng.standard_normal(n)
    P = 2.0
    s = orc.OracleSpline(k, f, periodic=[True], ranges=[(0.0, P)])
    x = np.arange(n + 1) * P / n
    sp = si.make_interp_spline(x, np.r_[f, f[0]], k=k, bc_type="periodic")
    q = rng.uniform(-1.0, 3.0, 60)
    assert np.allclose(s.eval(q[:, None])[0], sp(np.mod(q, P)), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [0, 2, 4, 6])
def test_oracle_periodic_even_dense_solve(orc, k):
    """Even order, periodic: build the cyclic collocation matrix with scipy's
    design matrix on the oracle's knots and solve densely."""
    n = 16
    rng = np.random.default_rng(300 + k)
    f = rng.standard_normal(n)
    s = orc.OracleSpline(k, f, periodic=[True], ranges=[(0.0, 1.0)])
    t = s.knots(0)
    x = np.arange(n) / n
    dm = si.BSpline.design_matrix(x, t, k, extrapolate=True).toarray()
    a = np.zeros((n, n))
    for j in range(dm.shape[1]):
        a[:, j % n] += dm[:, j]
    # the reference stores c at positions shifted by k/2 (RHS shift
    # InterpolationTemplate.hpp:452-462) -> the table index is the basis
    # index modulo n
    c = np.linalg.solve(a, f)
    assert np.allclose(s.control_points(), c, rtol=1e-12, atol=1e-12)
    assert np.allclose(s.eval(x[:, None])[0], f, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7])
def test_oracle_polynomial_reproduction_3d(orc, k):
    """Clamped data sampled from a polynomial of degree <= k is reproduced
    exactly (to round-off), values and partial derivatives."""
    rng = np.random.default_rng(400 + k)
    n = (9, 8, 7) if k < 5 else (10, 9, 8)
    ax = [np.linspace(-1.0, 2.0, n[0]), np.linspace(0.0, 1.0, n[1]), np.linspace(1.0, 3.0, n[2])]
    c = rng.standard_normal(3)

    def P(x, y, z):
        return c[0] * x ** k + c[1] * (x * y) ** (k // 2) * z ** (k - 2 * (k // 2)) + c[2] * y * z ** (k - 1)

    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    s = orc.OracleSpline(k, P(X, Y, Z), ranges=[(-1.0, 2.0), (0.0, 1.0), (1.0, 3.0)])
    q = np.stack([rng.uniform(-1, 2, 40), rng.uniform(0, 1, 40), rng.uniform(1, 3, 40)], 1)
    got = s.eval(q)[0]
    assert np.allclose(got, P(q[:, 0], q[:, 1], q[:, 2]), rtol=1e-11, atol=1e-11)
    # d/dz analytically
    dz = c[1] * (q[:, 0] * q[:, 1]) ** (k // 2) * (k % 2) + c[2] * q[:, 1] * (k - 1) * q[:, 2] ** max(k - 2, 0)
    assert np.allclose(s.eval(q, [[0, 0, 1]])[0], dz, rtol=1e-9, atol=1e-9)


def test_oracle_tensor_matches_ndbspline(orc):
    """3D clamped cubic, non-uniform grid spacing through uniform knots on
    each axis: the oracle's coefficients used by scipy's NdBSpline give the
    oracle's values and mixed partials."""
    rng = np.random.default_rng(9)
    f = rng.standard_normal((7, 9, 6))
    s = orc.OracleSpline(3, f, periodic=[False, False, False], ranges=[(0, 1), (0, 2), (-1, 1)])
    t = [s.knots(d) for d in range(3)]
    nd = si.NdBSpline(tuple(t), s.control_points(), 3)
    q = np.stack([rng.uniform(0, 1, 30), rng.uniform(0, 2, 30), rng.uniform(-1, 1, 30)], 1)
    assert np.allclose(s.eval(q)[0], nd(q), rtol=1e-12, atol=1e-12)
    for r in ([1, 0, 0], [0, 2, 1], [1, 1, 1], [3, 0, 2]):
        assert np.allclose(s.eval(q, [r])[0], nd(q, nu=r), rtol=1e-9, atol=1e-9)
    # data are interpolated at the grid nodes
    g = np.stack(np.meshgrid(np.linspace(0, 1, 7), np.linspace(0, 2, 9), np.linspace(-1, 1, 6),
                             indexing="ij"), -1).reshape(-1, 3)
    assert np.allclose(s.eval(g)[0], f.ravel(), rtol=1e-12, atol=1e-12)


def test_oracle_periodic_invariance(orc):
    rng = np.random.default_rng(12)
    f = rng.standard_normal((10, 9))
    for k in (2, 3, 4):
        s = orc.OracleSpline(k, f, periodic=[True, False], ranges=[(0.0, 5.0), (0.0, 1.0)])
        q = np.stack([rng.uniform(0, 5, 20), rng.uniform(0, 1, 20)], 1)
        v = s.eval(q, [[0, 0], [1, 0], [0, 1]])
        for shift in (-5.0, 5.0, 15.0):
            q2 = q.copy()
            q2[:, 0] += shift
            assert np.allclose(s.eval(q2, [[0, 0], [1, 0], [0, 1]]), v, rtol=1e-11, atol=1e-11)


def test_oracle_derivative_above_order_is_zero(orc):
    f = np.random.default_rng(1).standard_normal((6, 6))
    s = orc.OracleSpline(2, f, ranges=[(0, 1), (0, 1)])
    assert np.all(s.eval([[0.3, 0.4]], [[3, 0], [0, 5]]) == 0.0)


@pytest.mark.parametrize("D,k", [(4, 3), (5, 2), (5, 3), (6, 1), (6, 3)])
def test_oracle_separable_matches_1d_products(orc, D, k):
    """Dimensions without an instantiated reference (5, 6; 4 also against
    the compiled reference): interpolating a separable product
    f(x) = prod_d g_d(x_d) on a tensor grid gives the product of the 1D
    interpolants of the g_d (the tensor-product solve of a rank-one right-
    hand side is rank one), and a partial along d replaces factor d by its
    derivative.  The 1D oracle is pinned against the reference, so a wrong
    axis, stride or periodic shift in the D-dimensional oracle fails here."""
    rng = np.random.default_rng(10 * D + k)
    n = [int(rng.integers(max(k + 2, 5), 9)) for _ in range(D)]
    per = [bool(d % 2 == 0) for d in range(D)]
    ranges = [(float(-1 - d * 0.1), float(1 + d * 0.2)) for d in range(D)]
    gs = [rng.standard_normal(n[d]) for d in range(D)]
    f = gs[0]
    for d in range(1, D):
        f = np.multiply.outer(f, gs[d])
    s = orc.OracleSpline(k, f, periodic=per, ranges=ranges)
    ones = [orc.OracleSpline(k, gs[d], periodic=[per[d]], ranges=[ranges[d]]) for d in range(D)]
    q = np.stack([rng.uniform(ranges[d][0], ranges[d][1], 60) for d in range(D)], axis=1)
    orders = [[0] * D] + [list(e) for e in np.eye(D, dtype=int)]
    got = s.eval(q, orders)
    vals = [ones[d].eval(q[:, [d]], [[0]])[0] for d in range(D)]
    ders = [ones[d].eval(q[:, [d]], [[1]])[0] for d in range(D)]
    want = [np.prod(vals, axis=0)]
    for d in range(D):
        want.append(np.prod([ders[e] if e == d else vals[e] for e in range(D)], axis=0))
    for i in range(len(orders)):
        assert rel_err(got[i], want[i]) < 1e-12, (i, rel_err(got[i], want[i]))
