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
//    dot product with the tensor-product basis.
#include "host_eval.hpp"

#include <cmath>
#include <cstddef>

namespace intp_amd {

namespace {

constexpr int kMaxK = 7;
constexpr int kMaxStencil = (kMaxK + 1) * (kMaxK + 1) * (kMaxK + 1);

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

double eval_point(const HostSpline& s, const double* xin, const int32_t* r) {
    const int D = s.dim, k = s.order, w = k + 1;
    for (int d = 0; d < D; ++d)
        if (r[d] > k) return 0.0;
    double x[3];
    int64_t mu[3], stride[3];
    double basis[3][kMaxK + 1];
    int64_t st = 1;
    for (int d = D - 1; d >= 0; --d) {
        stride[d] = st;
        st *= s.dims[d]->n;
    }
    for (int d = 0; d < D; ++d) {
        x[d] = xin[d];
        mu[d] = span_of(*s.dims[d], k, x[d]);
        basis_values(s.dims[d]->knots.data(), mu[d], x[d], k, k - r[d], basis[d]);
    }
    // gather the local (k+1)^D control points, C order
    double c[kMaxStencil];
    int count = 1;
    for (int d = 0; d < D; ++d) count *= w;
    for (int l = 0; l < count; ++l) {
        int64_t at = 0;
        int rem = l;
        for (int d = D - 1; d >= 0; --d) {
            const int m = rem % w;
            rem /= w;
            at += stencil_index(*s.dims[d], mu[d] - k, m) * stride[d];
        }
        c[l] = s.ctrl[at];
    }
    // difference along each differentiated dimension
    for (int d = 0; d < D; ++d) {
        if (r[d] == 0) continue;
        const double* t = s.dims[d]->knots.data();
        int lstride = 1;
        for (int e = d + 1; e < D; ++e) lstride *= w;
        const int fibers = count / w;
        for (int f = 0; f < fibers; ++f) {
            const int outer = f / lstride, inner = f % lstride;
            double* v = c + outer * lstride * w + inner;
            for (int kk = k; kk > k - r[d]; --kk)
                for (int j = kk; j > 0; --j) {
                    const int at = k + j - kk;
                    v[at * lstride] = static_cast<double>(kk) * (v[at * lstride] - v[(at - 1) * lstride]) /
                                      (t[mu[d] + j] - t[mu[d] + j - kk]);
                }
        }
    }
    double val = 0.0;
    for (int l = 0; l < count; ++l) {
        double coef = 1.0;
        int rem = l;
        for (int d = D - 1; d >= 0; --d) {
            coef *= basis[d][rem % w];
            rem /= w;
        }
        val += coef * c[l];
    }
    return val;
}

}  // namespace

int64_t host_eval_points(const HostSpline& s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                         int nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, bool bounds) {
    const int D = s.dim;
    for (int64_t i = 0; i < npts; ++i) {
        double x[3];
        for (int d = 0; d < D; ++d) x[d] = coords[d * cs_dim + i * cs_pt];
        if (bounds)
            for (int d = 0; d < D; ++d)
                if (outside(*s.dims[d], x[d])) return i;
        for (int o = 0; o < nout; ++o) out[o * os_out + i * os_pt] = eval_point(s, x, orders + o * D);
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
    int64_t stride[3];
    int64_t st = 1;
    for (int d = D - 1; d >= 0; --d) {
        stride[d] = st;
        st *= s.dims[d]->n;
    }
    int count = 1;
    for (int d = 0; d < D; ++d) count *= w;
    for (int64_t i = 0; i < npts; ++i) {
        double v = 0.0;
        for (int l = 0; l < count; ++l) {
            int64_t at = 0;
            double coef = 1.0;
            int rem = l;
            for (int d = D - 1; d >= 0; --d) {
                const int m = rem % w;
                rem /= w;
                at += stencil_index(*s.dims[d], first[i * D + d], m) * stride[d];
                coef *= weights[(i * D + d) * w + m];
            }
            v += coef * s.ctrl[at];
        }
        out[i * os_pt] = v;
    }
}

}  // namespace intp_amd
