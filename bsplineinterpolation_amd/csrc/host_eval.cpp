// host_eval.cpp -- per-point evaluation on the host, the reference's
// algorithm restated for a host copy of a device-built table:
//  * periodic coordinates are wrapped into [lo, lo + period) as
//    BSpline.hpp:130-136 does;
//  * the knot span is the unique mu in [order, nt - order - 2] with
//    t[mu] <= x < t[mu + 1], clamped at the ends (BSpline.hpp:125-157),
//    found from the uniform-grid hint (Interpolation.hpp:277-295) or by
//    binary search;
//  * basis values by Cox-de Boor, a reduced-order basis right-aligned for
//    derivatives (BSpline.hpp:83-111);
//  * derivatives by differencing the gathered (order + 1)^D control points
//    along each differentiated dimension (BSpline.hpp:489-520), then one
//    dot product with the tensor-product basis;
//  * the products and sums in the reference's own order, so that values,
//    derivatives and proxy results equal the reference's bit for bit.
// Dimension and order are template arguments (a table of instantiations):
// the stencil loops unroll and no index is split by division.  Splines of
// 4 to 6 dimensions take a runtime-dimension form of the same loops, with
// the same operation order.
#include "host_eval.hpp"

#include <array>
#include <cmath>
#include <cstddef>
#include <utility>
#include <vector>

namespace intp_amd {

namespace {

constexpr int kMaxK = 7;

// Wraps a periodic coordinate and returns its knot span.
int64_t span_of(const DimSetup& s, int k, double& x) {
    if (s.periodic) {
        const double period = s.hi - s.lo;
        x = s.lo + std::fmod(x - s.lo, period) + (x < s.lo ? period : 0.0);
    }
    const double* t = s.knots.data();
    const int64_t last = static_cast<int64_t>(s.knots.size()) - k - 2;
    if (s.uniform && s.dx > 0.0) {
        const double v = (x - s.lo) / s.dx - (s.periodic ? 1.0 : 0.5 * static_cast<double>(k + 1));
        if (v >= 0.0 && v < static_cast<double>(last)) {
            int64_t mu = static_cast<int64_t>(std::ceil(v)) + k;
            if (mu > last) mu = last;
            if (t[mu] <= x && x < t[mu + 1]) return mu;
            if (mu > k && t[mu - 1] <= x && x < t[mu]) return mu - 1;
            if (mu < last && t[mu + 1] <= x && x < t[mu + 2]) return mu + 1;
        }
    }
    int64_t lo = k + 1, hi = last + 1;  // upper_bound over t[k+1 .. last]
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (x < t[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo - 1;
}

bool outside(const DimSetup& s, double x) { return !s.periodic && (x < s.lo || x > s.hi); }

// Table index of stencil element m along a dimension.
inline int64_t stencil_index(const DimSetup& s, int64_t first, int m) {
    int64_t id = first + m;
    if (s.periodic && id >= s.n) id -= s.n;
    return id;
}

// Per dimension: wrap, span, the K + 1 basis values of order K - r
// (right-aligned) and the table offsets of the stencil's elements.
template <int K>
inline int64_t dimension_terms(const DimSetup& ds, int64_t stride, double x, int r, double* b, int64_t* off) {
    const int64_t mu = span_of(ds, K, x);
    basis_values(ds.knots.data(), mu, x, K, K - r, b);
    const int64_t first = mu - K;
    for (int m = 0; m <= K; ++m) off[m] = stencil_index(ds, first, m) * stride;
    return mu;
}

// All dimensions; returns the spans in mu.
template <int D, int K>
inline void all_terms(const HostSpline& s, const double* x, const int32_t* r, double (&b)[3][K + 1],
                      int64_t (&off)[3][K + 1], int64_t (&mu)[3]) {
    int64_t stride = 1;
    for (int d = D - 1; d >= 0; --d) {
        mu[d] = dimension_terms<K>(*s.dims[d], stride, x[d], r ? r[d] : 0, b[d], off[d]);
        stride *= s.dims[d]->n;
    }
}

// Value, summed as the reference's operator() does with its cell layout
// (BSpline.hpp:305-335): stencil elements with dimension 0 varying
// fastest, each term ((c * b_0) * b_1) * b_2, added in that order.
template <int D, int K>
double value_at(const HostSpline& s, const double* x) {
    constexpr int W = K + 1;
    double b[3][W];
    int64_t off[3][W], mu[3];
    all_terms<D, K>(s, x, nullptr, b, off, mu);
    const double* c = s.ctrl;
    double v = 0.0;
    if constexpr (D == 1) {
        for (int i0 = 0; i0 < W; ++i0) {
            double coef = c[off[0][i0]];
            coef *= b[0][i0];
            v += coef;
        }
    } else if constexpr (D == 2) {
        for (int i1 = 0; i1 < W; ++i1)
            for (int i0 = 0; i0 < W; ++i0) {
                double coef = c[off[0][i0] + off[1][i1]];
                coef *= b[0][i0];
                coef *= b[1][i1];
                v += coef;
            }
    } else {
        for (int i2 = 0; i2 < W; ++i2)
            for (int i1 = 0; i1 < W; ++i1)
                for (int i0 = 0; i0 < W; ++i0) {
                    double coef = c[off[0][i0] + off[1][i1] + off[2][i2]];
                    coef *= b[0][i0];
                    coef *= b[1][i1];
                    coef *= b[2][i2];
                    v += coef;
                }
    }
    return v;
}

// Partial derivative, as the reference's derivative_at (BSpline.hpp:393-531):
// the local (K+1)^D control points in C order, differenced along each
// differentiated dimension, then sum over C order of (b_0 b_1 b_2) * c.
// A first partial of a refined spline takes its differences from the
// gradient table (s.diff) instead of differencing the FP32 points.
template <int D, int K>
double derivative_at(const HostSpline& s, const double* x, const int32_t* r) {
    constexpr int W = K + 1;
    constexpr int N = D == 1 ? W : D == 2 ? W * W : W * W * W;
    for (int d = 0; d < D; ++d)
        if (r[d] > K) return 0.0;
    int first = -1;  // dimension of a first partial read from s.diff
    if (s.diff) {
        int sum = 0;
        for (int d = 0; d < D; ++d) sum += r[d];
        for (int d = 0; d < D && sum == 1; ++d)
            if (r[d] == 1) first = d;
    }
    double b[3][W];
    int64_t off[3][W], mu[3];
    all_terms<D, K>(s, x, r, b, off, mu);
    const double* src = first >= 0 ? s.diff + first * s.total : s.ctrl;
    double c[N];
    int l = 0;
    if constexpr (D == 1) {
        for (int i0 = 0; i0 < W; ++i0) c[l++] = src[off[0][i0]];
    } else if constexpr (D == 2) {
        for (int i0 = 0; i0 < W; ++i0)
            for (int i1 = 0; i1 < W; ++i1) c[l++] = src[off[0][i0] + off[1][i1]];
    } else {
        for (int i0 = 0; i0 < W; ++i0)
            for (int i1 = 0; i1 < W; ++i1)
                for (int i2 = 0; i2 < W; ++i2) c[l++] = src[off[0][i0] + off[1][i1] + off[2][i2]];
    }
    for (int d = 0; d < D; ++d) {
        if (r[d] == 0) continue;
        const double* t = s.dims[d]->knots.data();
        int lstride = 1;
        for (int e = d + 1; e < D; ++e) lstride *= W;
        constexpr int fibers = N / W;
        for (int f = 0; f < fibers; ++f) {
            const int outer = f / lstride, inner = f % lstride;
            double* v = c + outer * lstride * W + inner;
            for (int kk = K; kk > K - r[d]; --kk)
                for (int j = kk; j > 0; --j) {
                    const int at = K + j - kk;
                    // gathered differences: v[at] already holds c[at] - c[at - 1]
                    const double dv = d == first ? v[at * lstride] : v[at * lstride] - v[(at - 1) * lstride];
                    v[at * lstride] = static_cast<double>(kk) * dv / (t[mu[d] + j] - t[mu[d] + j - kk]);
                }
        }
    }
    double v = 0.0;
    l = 0;
    if constexpr (D == 1) {
        for (int i0 = 0; i0 < W; ++i0) v += b[0][i0] * c[l++];
    } else if constexpr (D == 2) {
        for (int i0 = 0; i0 < W; ++i0)
            for (int i1 = 0; i1 < W; ++i1) {
                double coef = b[0][i0];
                coef *= b[1][i1];
                v += coef * c[l++];
            }
    } else {
        for (int i0 = 0; i0 < W; ++i0)
            for (int i1 = 0; i1 < W; ++i1)
                for (int i2 = 0; i2 < W; ++i2) {
                    double coef = b[0][i0];
                    coef *= b[1][i1];
                    coef *= b[2][i2];
                    v += coef * c[l++];
                }
    }
    return v;
}

template <int D, int K>
double eval_point_t(const HostSpline& s, const double* x, const int32_t* r) {
    for (int d = 0; d < D; ++d)
        if (r[d] != 0) return derivative_at<D, K>(s, x, r);
    return value_at<D, K>(s, x);
}

using PointFn = double (*)(const HostSpline&, const double*, const int32_t*);

template <int D, int... K>
constexpr std::array<PointFn, sizeof...(K)> point_fns(std::integer_sequence<int, K...>) {
    return {&eval_point_t<D, K>...};
}

const std::array<std::array<PointFn, kMaxK + 1>, 3> kPointFns = {
    point_fns<1>(std::make_integer_sequence<int, kMaxK + 1>{}),
    point_fns<2>(std::make_integer_sequence<int, kMaxK + 1>{}),
    point_fns<3>(std::make_integer_sequence<int, kMaxK + 1>{})};

// Proxy application, summed like value_at (the reference's operator()
// order), so that a host proxy gives exactly the value the scalar call
// gives.  (The reference's own pre_calc_coef closure, BSpline.hpp:270-292,
// multiplies the weights first and may differ from its operator() in the
// last bit.)
template <int D, int K>
double apply_weights(const HostSpline& s, const int32_t* first, const double* w) {
    constexpr int W = K + 1;
    int64_t off[3][W];
    int64_t stride = 1;
    for (int d = D - 1; d >= 0; --d) {
        for (int m = 0; m < W; ++m) off[d][m] = stencil_index(*s.dims[d], first[d], m) * stride;
        stride *= s.dims[d]->n;
    }
    const double* c = s.ctrl;
    double v = 0.0;
    if constexpr (D == 1) {
        for (int i0 = 0; i0 < W; ++i0) {
            double coef = c[off[0][i0]];
            coef *= w[i0];
            v += coef;
        }
    } else if constexpr (D == 2) {
        for (int i1 = 0; i1 < W; ++i1)
            for (int i0 = 0; i0 < W; ++i0) {
                double coef = c[off[0][i0] + off[1][i1]];
                coef *= w[i0];
                coef *= w[W + i1];
                v += coef;
            }
    } else {
        for (int i2 = 0; i2 < W; ++i2)
            for (int i1 = 0; i1 < W; ++i1)
                for (int i0 = 0; i0 < W; ++i0) {
                    double coef = c[off[0][i0] + off[1][i1] + off[2][i2]];
                    coef *= w[i0];
                    coef *= w[W + i1];
                    coef *= w[2 * W + i2];
                    v += coef;
                }
    }
    return v;
}

using ApplyFn = double (*)(const HostSpline&, const int32_t*, const double*);

template <int D, int... K>
constexpr std::array<ApplyFn, sizeof...(K)> apply_fns(std::integer_sequence<int, K...>) {
    return {&apply_weights<D, K>...};
}

const std::array<std::array<ApplyFn, kMaxK + 1>, 3> kApplyFns = {
    apply_fns<1>(std::make_integer_sequence<int, kMaxK + 1>{}),
    apply_fns<2>(std::make_integer_sequence<int, kMaxK + 1>{}),
    apply_fns<3>(std::make_integer_sequence<int, kMaxK + 1>{})};

// ---- 4 to 6 dimensions: the same loops, nested by recursion over the
// dimensions (D a template argument); the offsets of the outer dimensions
// are summed once per level ----

template <int D, int K>
void all_terms_nd(const HostSpline& s, const double* x, const int32_t* r, double (&b)[kHostMaxDim][K + 1],
                  int64_t (&off)[kHostMaxDim][K + 1], int64_t (&mu)[kHostMaxDim]) {
    int64_t stride = 1;
    for (int d = D - 1; d >= 0; --d) {
        mu[d] = dimension_terms<K>(*s.dims[d], stride, x[d], r ? r[d] : 0, b[d], off[d]);
        stride *= s.dims[d]->n;
    }
}

// Cell order, dimension 0 innermost: each term ((c * w_0) * w_1) ... added
// in turn (value_at's order).
template <int D, int K, int L>
inline void cell_sum(const double* c, const double (&w)[kHostMaxDim][K + 1], const int64_t (&off)[kHostMaxDim][K + 1],
                     int64_t base, int (&a)[kHostMaxDim], double& v) {
    for (int i = 0; i <= K; ++i) {
        a[L] = i;
        if constexpr (L == 0) {
            double coef = c[base + off[0][i]];
            coef *= w[0][i];
            for (int d = 1; d < D; ++d) coef *= w[d][a[d]];
            v += coef;
        } else {
            cell_sum<D, K, L - 1>(c, w, off, base + off[L][i], a, v);
        }
    }
}

// C order, dimension D-1 innermost: the stencil's elements into out.
template <int D, int K, int L>
inline void gather_c_order(const double* c, const int64_t (&off)[kHostMaxDim][K + 1], int64_t base, double*& out) {
    for (int i = 0; i <= K; ++i) {
        if constexpr (L == D - 1)
            *out++ = c[base + off[L][i]];
        else
            gather_c_order<D, K, L + 1>(c, off, base + off[L][i], out);
    }
}

// C-order sum of (b_0 b_1 ...) * c: the product b_0 b_1 .. b_L is formed
// left to right as in derivative_at, its prefix shared by the terms below
// a level.
template <int D, int K, int L>
inline void c_order_sum(const double (&b)[kHostMaxDim][K + 1], double prefix, const double*& c, double& v) {
    for (int i = 0; i <= K; ++i) {
        const double q = L == 0 ? b[0][i] : prefix * b[L][i];
        if constexpr (L == D - 1)
            v += q * *c++;
        else
            c_order_sum<D, K, L + 1>(b, q, c, v);
    }
}

template <int D, int K>
double value_at_nd(const HostSpline& s, const double* x) {
    double b[kHostMaxDim][K + 1];
    int64_t off[kHostMaxDim][K + 1], mu[kHostMaxDim];
    all_terms_nd<D, K>(s, x, nullptr, b, off, mu);
    int a[kHostMaxDim] = {};
    double v = 0.0;
    cell_sum<D, K, D - 1>(s.ctrl, b, off, 0, a, v);
    return v;
}

template <int D, int K>
double derivative_at_nd(const HostSpline& s, const double* x, const int32_t* r) {
    constexpr int W = K + 1;
    int64_t total = 1;
    for (int d = 0; d < D; ++d) total *= W;
    for (int d = 0; d < D; ++d)
        if (r[d] > K) return 0.0;
    double b[kHostMaxDim][K + 1];
    int64_t off[kHostMaxDim][K + 1], mu[kHostMaxDim];
    all_terms_nd<D, K>(s, x, r, b, off, mu);
    thread_local std::vector<double> buf;
    buf.resize(static_cast<size_t>(total));
    double* c = buf.data();
    {
        double* out = c;
        gather_c_order<D, K, 0>(s.ctrl, off, 0, out);
    }
    for (int d = 0; d < D; ++d) {
        if (r[d] == 0) continue;
        const double* t = s.dims[d]->knots.data();
        int64_t lstride = 1;
        for (int e = d + 1; e < D; ++e) lstride *= W;
        const int64_t fibers = total / W;
        for (int64_t f = 0; f < fibers; ++f) {
            const int64_t outer = f / lstride, inner = f % lstride;
            double* v = c + outer * lstride * W + inner;
            for (int kk = K; kk > K - r[d]; --kk)
                for (int j = kk; j > 0; --j) {
                    const int at = K + j - kk;
                    v[at * lstride] = static_cast<double>(kk) * (v[at * lstride] - v[(at - 1) * lstride]) /
                                      (t[mu[d] + j] - t[mu[d] + j - kk]);
                }
        }
    }
    double v = 0.0;
    const double* in = c;
    c_order_sum<D, K, 0>(b, 1.0, in, v);
    return v;
}

template <int D, int K>
double eval_point_nd(const HostSpline& s, const double* x, const int32_t* r) {
    for (int d = 0; d < D; ++d)
        if (r[d] != 0) return derivative_at_nd<D, K>(s, x, r);
    return value_at_nd<D, K>(s, x);
}

// apply_weights for D >= 4 (value_at_nd's order)
template <int D, int K>
double apply_weights_nd(const HostSpline& s, const int32_t* first, const double* w) {
    constexpr int W = K + 1;
    int64_t off[kHostMaxDim][W];
    double wl[kHostMaxDim][W];
    int64_t stride = 1;
    for (int d = D - 1; d >= 0; --d) {
        for (int m = 0; m < W; ++m) {
            off[d][m] = stencil_index(*s.dims[d], first[d], m) * stride;
            wl[d][m] = w[d * W + m];
        }
        stride *= s.dims[d]->n;
    }
    int a[kHostMaxDim] = {};
    double v = 0.0;
    cell_sum<D, K, D - 1>(s.ctrl, wl, off, 0, a, v);
    return v;
}

template <int D, int... K>
constexpr std::array<PointFn, sizeof...(K)> point_fns_nd(std::integer_sequence<int, K...>) {
    return {&eval_point_nd<D, K>...};
}
template <int D, int... K>
constexpr std::array<ApplyFn, sizeof...(K)> apply_fns_nd(std::integer_sequence<int, K...>) {
    return {&apply_weights_nd<D, K>...};
}
const std::array<std::array<PointFn, kMaxK + 1>, 3> kPointFnsNd = {
    point_fns_nd<4>(std::make_integer_sequence<int, kMaxK + 1>{}),
    point_fns_nd<5>(std::make_integer_sequence<int, kMaxK + 1>{}),
    point_fns_nd<6>(std::make_integer_sequence<int, kMaxK + 1>{})};
const std::array<std::array<ApplyFn, kMaxK + 1>, 3> kApplyFnsNd = {
    apply_fns_nd<4>(std::make_integer_sequence<int, kMaxK + 1>{}),
    apply_fns_nd<5>(std::make_integer_sequence<int, kMaxK + 1>{}),
    apply_fns_nd<6>(std::make_integer_sequence<int, kMaxK + 1>{})};

PointFn point_fn(int dim, int order) { return dim <= 3 ? kPointFns[dim - 1][order] : kPointFnsNd[dim - 4][order]; }
ApplyFn apply_fn(int dim, int order) { return dim <= 3 ? kApplyFns[dim - 1][order] : kApplyFnsNd[dim - 4][order]; }

}  // namespace

int64_t host_eval_points(const HostSpline& s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                         int nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, bool bounds) {
    const int D = s.dim;
    const PointFn fn = point_fn(D, s.order);
    for (int64_t i = 0; i < npts; ++i) {
        double x[kHostMaxDim];
        for (int d = 0; d < D; ++d) x[d] = coords[d * cs_dim + i * cs_pt];
        if (bounds)
            for (int d = 0; d < D; ++d)
                if (outside(*s.dims[d], x[d])) return i;
        for (int o = 0; o < nout; ++o) out[o * os_out + i * os_pt] = fn(s, x, orders + o * D);
    }
    return -1;
}

int64_t host_proxy_weights(const HostSpline& s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                           bool bounds, int32_t* first, double* weights) {
    const int D = s.dim, k = s.order, w = k + 1;
    for (int64_t i = 0; i < npts; ++i) {
        for (int d = 0; d < D; ++d) {
            double x = coords[d * cs_dim + i * cs_pt];
            if (bounds && outside(*s.dims[d], x)) return i;
            const int64_t mu = span_of(*s.dims[d], k, x);
            first[i * D + d] = static_cast<int32_t>(mu - k);
            basis_values(s.dims[d]->knots.data(), mu, x, k, k, weights + (i * D + d) * w);
        }
    }
    return -1;
}

void host_proxy_apply(const HostSpline& s, int64_t npts, const int32_t* first, const double* weights, double* out,
                      int64_t os_pt) {
    const int D = s.dim, w = s.order + 1;
    const ApplyFn fn = apply_fn(D, s.order);
    for (int64_t i = 0; i < npts; ++i) out[i * os_pt] = fn(s, first + i * D, weights + i * D * w);
}

}  // namespace intp_amd
