// setup.cpp -- knot vectors, collocation matrices and their band LU.
//
// Follows the reference's construction so the factors match it bit for bit
// (compiled without FMA contraction): knots Interpolation.hpp:322-464,
// range BSpline.hpp:224-226, collocation rows InterpolationTemplate.hpp:
// 254-426, no-pivot LU BandLU.hpp:103-118,159-213.  The LU here runs over
// the fill pattern of a cyclic band matrix (band + right columns + bottom
// rows); every entry receives the reference's updates in the reference's
// order (pivot step ascending), which is what fixes its rounding.
#include "setup.hpp"

#include <algorithm>
#include <cmath>

namespace intp_amd {

namespace {

// The fill pattern of a (cyclic) band matrix and its LU.
class PatternMatrix {
   public:
    PatternMatrix(int64_t n, int p, int q, bool cyclic)
        : n_(n), p_(p), q_(q), w_(p + q + 1), cyclic_(cyclic),
          band_(static_cast<size_t>(n * w_), 0.0),
          right_(cyclic ? static_cast<size_t>(n * p) : 0, 0.0),
          bottom_(cyclic ? static_cast<size_t>(q * n) : 0, 0.0) {}

    bool in_band(int64_t i, int64_t j) const { return j - i <= q_ && i - j <= p_; }

    double& at(int64_t i, int64_t j) {
        if (in_band(i, j)) return band_[static_cast<size_t>(i * w_ + (j - i + p_))];
        if (cyclic_ && j >= n_ - p_) return right_[static_cast<size_t>(i * p_ + (j - (n_ - p_)))];
        // cyclic bottom rows
        return bottom_[static_cast<size_t>((i - (n_ - q_)) * n_ + j)];
    }
    bool stored(int64_t i, int64_t j) const {
        if (i < 0 || j < 0 || i >= n_ || j >= n_) return false;
        if (in_band(i, j)) return true;
        return cyclic_ && (j >= n_ - p_ || i >= n_ - q_);
    }

    // Right-looking LU without pivoting over the fill pattern.
    void factor() {
        for (int64_t k = 0; k + 1 < n_; ++k) {
            const int64_t r0 = k + 1, r1 = std::min<int64_t>(k + p_, n_ - 1);
            const int64_t b0 = cyclic_ ? std::max<int64_t>(n_ - q_, k + p_ + 1) : n_;
            const int64_t c0 = k + 1, c1 = std::min<int64_t>(k + q_, n_ - 1);
            const int64_t s0 = cyclic_ ? std::max<int64_t>(n_ - p_, k + q_ + 1) : n_;
            const double pivot = at(k, k);
            for (int64_t i = r0; i <= r1; ++i) at(i, k) /= pivot;
            for (int64_t i = b0; i < n_; ++i) at(i, k) /= pivot;
            auto update_row = [&](int64_t i) {
                const double lik = at(i, k);
                for (int64_t j = c0; j <= c1; ++j) at(i, j) -= lik * at(k, j);
                for (int64_t j = s0; j < n_; ++j) at(i, j) -= lik * at(k, j);
            };
            for (int64_t i = r0; i <= r1; ++i) update_row(i);
            for (int64_t i = b0; i < n_; ++i) update_row(i);
        }
    }

    void export_factors(LineFactors& f) {
        f.n = n_;
        f.p = p_;
        f.q = q_;
        f.cyclic = cyclic_;
        f.lower.assign(static_cast<size_t>(n_ * p_), 0.0);
        f.upper.assign(static_cast<size_t>(n_ * q_), 0.0);
        f.diag.assign(static_cast<size_t>(n_), 0.0);
        for (int64_t i = 0; i < n_; ++i) {
            f.diag[static_cast<size_t>(i)] = at(i, i);
            for (int m = 1; m <= p_; ++m)
                if (i - m >= 0) f.lower[static_cast<size_t>(i * p_ + m - 1)] = at(i, i - m);
            for (int m = 1; m <= q_; ++m)
                if (i + m < n_) f.upper[static_cast<size_t>(i * q_ + m - 1)] = at(i, i + m);
        }
        if (cyclic_) {
            f.bottom.assign(static_cast<size_t>(q_ * n_), 0.0);
            f.right.assign(static_cast<size_t>(p_ * n_), 0.0);
            for (int r = 0; r < q_; ++r) {
                const int64_t i = n_ - q_ + r;
                for (int64_t j = 0; j < i - p_; ++j) f.bottom[static_cast<size_t>(r * n_ + j)] = at(i, j);
            }
            for (int c = 0; c < p_; ++c) {
                const int64_t j = n_ - p_ + c;
                for (int64_t i = 0; i < j - q_; ++i) f.right[static_cast<size_t>(c * n_ + i)] = at(i, j);
            }
        }
    }

   private:
    int64_t n_;
    int p_, q_, w_;
    bool cyclic_;
    std::vector<double> band_, right_, bottom_;
};

// Span containing x among t[order .. last]: the unique mu with
// t[mu] <= x < t[mu+1], clamped to [order, last] (BSpline.hpp:146-156).
int64_t find_span(const std::vector<double>& t, int order, double x, int64_t last) {
    int64_t lo = order + 1, hi = last + 1;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (x < t[static_cast<size_t>(mid)])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo - 1;
}

void fill_recip(int order, DimSetup& s) {
    const int64_t nt = static_cast<int64_t>(s.knots.size());
    const int T = order * (order + 1) / 2;
    s.recip.assign(static_cast<size_t>(nt * std::max(T, 1)), 0.0);
    const double* t = s.knots.data();
    for (int64_t mu = order; mu <= nt - order - 2; ++mu)
        for (int i = 1; i <= order; ++i)
            for (int a = 1; a <= i; ++a) {
                const double d = t[mu + a] - t[mu + a - i];
                s.recip[static_cast<size_t>(mu * T + i * (i - 1) / 2 + a - 1)] = d != 0.0 ? 1.0 / d : 0.0;
            }
}

// Power-basis coefficients (in u = (x - t[mu]) / dx) of the order+1 B-splines
// non-zero on span mu: the Cox-de Boor recursion run on polynomials.
void span_polynomials(const std::vector<double>& t, int k, int64_t mu, double dx, double* out) {
    const int w = k + 1;
    std::vector<double> tau(static_cast<size_t>(2 * k + 2), 0.0);  // tau[i + k] = knot mu+i, i in [-k, k+1]
    for (int i = -k; i <= k + 1; ++i) {
        const int64_t idx = mu + i;
        if (idx >= 0 && idx < static_cast<int64_t>(t.size()))
            tau[static_cast<size_t>(i + k)] = (t[static_cast<size_t>(idx)] - t[static_cast<size_t>(mu)]) / dx;
    }
    auto T = [&](int i) { return tau[static_cast<size_t>(i + k)]; };
    std::vector<double> b(static_cast<size_t>(w * w), 0.0);  // b[j*w + p]
    b[static_cast<size_t>(k * w)] = 1.0;
    for (int i = 1; i <= k; ++i) {
        const int first = k - i;
        for (int j = 0; j <= i; ++j) {
            const int idx = first + j;
            std::vector<double> nb(static_cast<size_t>(w), 0.0);
            if (j != 0) {  // b[idx] * (u - tau_{j-i}) / (tau_j - tau_{j-i})
                const double lo = T(j - i), den = T(j) - T(j - i);
                for (int p = 0; p < w; ++p) {
                    const double c = b[static_cast<size_t>(idx * w + p)] / den;
                    if (p + 1 < w) nb[static_cast<size_t>(p + 1)] += c;
                    nb[static_cast<size_t>(p)] -= c * lo;
                }
            }
            if (idx != k) {  // b[idx+1] * (tau_{j+1} - u) / (tau_{j+1} - tau_{j+1-i})
                const double hi = T(j + 1), den = T(j + 1) - T(j + 1 - i);
                for (int p = 0; p < w; ++p) {
                    const double c = b[static_cast<size_t>((idx + 1) * w + p)] / den;
                    nb[static_cast<size_t>(p)] += c * hi;
                    if (p + 1 < w) nb[static_cast<size_t>(p + 1)] -= c;
                }
            }
            for (int p = 0; p < w; ++p) b[static_cast<size_t>(idx * w + p)] = nb[static_cast<size_t>(p)];
        }
    }
    for (int q = 0; q < w * w; ++q) out[q] = b[static_cast<size_t>(q)];
}

void fill_uniform_classes(int k, DimSetup& s, double xmin, double xmax) {
    const int64_t nt = static_cast<int64_t>(s.knots.size());
    const int64_t last = nt - k - 2;
    const int64_t nspan = last - k + 1;
    s.xmin = xmin;
    s.xmax = xmax;
    s.half = 0.5 * static_cast<double>(nt - (s.n + (s.periodic ? 1 : 0)));
    if (s.periodic) {
        s.nleft = s.nright = 0;
    } else {
        // span mu is interior when knots mu-k+1 .. mu+k all lie in [k+1, nt-k-2]
        s.nleft = static_cast<int>(std::min<int64_t>(nspan, k));
        s.nright = static_cast<int>(std::min<int64_t>(nspan - s.nleft, k));
    }
    s.nclass = 1 + s.nleft + s.nright;
    const int w = k + 1;
    s.poly.assign(static_cast<size_t>(s.nclass * w * w), 0.0);
    const int64_t interior = s.periodic ? k : k + s.nleft;
    if (interior <= last - s.nright) span_polynomials(s.knots, k, interior, s.dx, &s.poly[0]);
    for (int c = 0; c < s.nleft; ++c)
        span_polynomials(s.knots, k, k + c, s.dx, &s.poly[static_cast<size_t>((1 + c) * w * w)]);
    for (int c = 0; c < s.nright; ++c)
        span_polynomials(s.knots, k, last - c, s.dx, &s.poly[static_cast<size_t>((1 + s.nleft + c) * w * w)]);
}

void set_range(int order, DimSetup& s) {
    const size_t nt = s.knots.size();
    s.lo = s.knots[static_cast<size_t>(order)];
    s.hi = s.knots[nt - static_cast<size_t>(order) - (2 - static_cast<size_t>(order % 2))];
}

}  // namespace

int64_t knot_count(int order, bool periodic, int64_t n) {
    const int64_t np = n + (periodic ? 1 : 0);
    return np + (periodic ? 2 * order + (1 - order % 2) : order + 1);
}

void basis_values(const double* t, int64_t mu, double x, int order, int spline_order, double* b) {
    for (int i = 0; i <= order; ++i) b[i] = 0.0;
    b[order] = 1.0;
    for (int i = 1; i <= spline_order; ++i) {
        const int first = order - i;
        for (int j = 0; j <= i; ++j) {
            const int64_t l = mu - (i - j), r = mu + j + 1;
            const double left = j == 0 ? 0.0 : b[first + j] * (x - t[l]) / (t[r - 1] - t[l]);
            const double right = first + j == order ? 0.0 : b[first + j + 1] * (t[r] - x) / (t[r] - t[l + 1]);
            b[first + j] = left + right;
        }
    }
}

std::string setup_dimension(int order, bool periodic, int64_t n, double xmin, double xmax,
                            const double* nodes, DimSetup& s) {
    const int k = order;
    if (n < std::max<int64_t>(2, k + 1)) return "too few values along a dimension for the spline order";
    const int bw = periodic ? k / 2 : (k == 0 ? 0 : k - 1);
    if (periodic && n < 2 * bw + 2) return "too few values along a periodic dimension for the spline order";
    s.periodic = periodic;
    s.n = n;
    const int64_t np = n + (periodic ? 1 : 0);
    const int64_t nt = knot_count(k, periodic, n);
    s.knots.assign(static_cast<size_t>(nt), 0.0);
    double* t = s.knots.data();
    if (!nodes) {
        // uniform knots: Interpolation.hpp:322-362
        if (!(xmax > xmin)) return "range maximum must exceed the minimum";
        s.uniform = true;
        s.dx = (xmax - xmin) / static_cast<double>(np - 1);
        const int64_t extra = nt - np;
        const double half = 0.5 * static_cast<double>(extra);
        for (int64_t i = 0; i < nt; ++i) t[i] = xmin;
        if (periodic) {
            for (int64_t i = 0; i < nt; ++i) t[i] = xmin + (static_cast<double>(i) - half) * s.dx;
        } else {
            for (int64_t i = k + 1; i < nt - k - 1; ++i) t[i] = xmin + (static_cast<double>(i) - half) * s.dx;
            for (int64_t i = nt - k - 1; i < nt; ++i) t[i] = xmax;
        }
        set_range(k, s);
        fill_uniform_classes(k, s, xmin, xmax);
    } else {
        // node-coordinate knots: Interpolation.hpp:365-447
        s.uniform = false;
        s.nodes.assign(nodes, nodes + np);
        for (int64_t i = 0; i + 1 < np; ++i)
            if (!(nodes[i + 1] > nodes[i])) return "Given coordinate is not monotonically increasing";
        if (periodic) {
            for (int64_t i = k + 1; i < k + np; ++i)
                t[i] = k % 2 == 0 ? 0.5 * (nodes[i - k - 1] + nodes[i - k]) : nodes[i - k];
            const double period = nodes[np - 1] - nodes[0];
            for (int64_t i = 0; i < k + 1; ++i) {
                t[i] = t[np + i - 1] - period;
                t[nt - i - 1] = t[nt - i - np] + period;
            }
        } else {
            if (k == 0) return "order 0 is not defined on a non-uniform dimension";
            for (int64_t i = 0; i <= k; ++i) t[i] = nodes[0];
            double window = 0.0;
            for (int64_t i = 1; i < k; ++i) window += nodes[i];
            for (int64_t i = k + 1; i < np; ++i) {
                window += nodes[i - 1];
                t[i] = window / static_cast<double>(k);
                window -= nodes[i - k];
            }
            for (int64_t i = np; i < np + k + 1; ++i) t[i] = nodes[np - 1];
        }
    }
    set_range(k, s);

    // collocation matrix: InterpolationTemplate.hpp:254-426
    PatternMatrix m(n, bw, bw, periodic);
    double bv[64];
    if (periodic && s.uniform)
        basis_values(t, k, t[k] + static_cast<double>(1 - k % 2) * s.dx * 0.5, k, k, bv);
    for (int64_t i = 0; i < n; ++i) {
        if (!periodic && (i == 0 || i == n - 1)) {
            m.at(i, i) = 1.0;
            continue;
        }
        int64_t mu;
        const bool internal = i > k / 2 && i < n - k / 2 - 1;
        if (s.uniform) {
            mu = periodic ? i + k : std::min<int64_t>(nt - k - 2, i > k / 2 ? i + (k + 1) / 2 : k);
            // clamped rows near the ends recompute the basis; interior rows
            // reuse the last one (the interior basis is translation
            // invariant on a uniform grid)
            if (!periodic && (mu <= 2 * k + 1 || mu >= nt - 2 * k - 2))
                basis_values(t, mu, s.lo + static_cast<double>(i) * s.dx, k, k, bv);
        } else {
            const double x = s.nodes[static_cast<size_t>(i)];
            if (periodic) {
                mu = i + k;
            } else {
                const int64_t last = std::min<int64_t>(nt - k - 1, i + k);
                const int64_t hint = i + 1;
                mu = (t[hint] <= x && t[hint + 1] > x) ? hint : find_span(s.knots, k, x, last);
            }
            basis_values(t, mu, x, k, k, bv);
        }
        const int64_t cnt = periodic ? (k | 1) : k == 1 ? 1 : (s.uniform && internal) ? (k | 1) : k + 1;
        for (int64_t j = 0; j < cnt; ++j) {
            const int64_t row = (i + (periodic ? bw : 0)) % n;
            const int64_t col = (mu - k + j) % n;
            if (!m.stored(row, col)) return "collocation entry outside the band pattern";
            m.at(row, col) = bv[j];
        }
    }
    // the unfactored matrix, row-wise, for FP64 residuals
    s.colloc_w = bw;
    s.colloc.assign(static_cast<size_t>(n * (2 * bw + 1)), 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int mm = -bw; mm <= bw; ++mm) {
            int64_t j = i + mm;
            if (periodic) j = ((j % n) + n) % n;
            if (j < 0 || j >= n || !m.stored(i, j)) continue;
            s.colloc[static_cast<size_t>(i * (2 * bw + 1) + (mm + bw))] = m.at(i, j);
        }
    m.factor();
    m.export_factors(s.lu);
    auto round32 = [](const std::vector<double>& v, std::vector<float>& out) {
        out.resize(v.size());
        for (size_t i = 0; i < v.size(); ++i) out[i] = static_cast<float>(v[i]);
    };
    round32(s.lu.lower, s.lower32);
    round32(s.lu.upper, s.upper32);
    round32(s.lu.diag, s.diag32);
    round32(s.lu.bottom, s.bottom32);
    round32(s.lu.right, s.right32);
    for (int64_t i = 0; i < n; ++i)
        if (!(std::fabs(s.lu.diag[static_cast<size_t>(i)]) > 0.0) || !std::isfinite(s.lu.diag[static_cast<size_t>(i)]))
            return "singular collocation matrix";
    fill_recip(k, s);
    return "";
}

std::string setup_from_knots(int order, bool periodic, int64_t n, const double* knots, int64_t nknots,
                             DimSetup& s) {
    const int64_t want = n + order + 1 + (periodic ? order : 0);
    if (nknots != want) return "Inconsistency between knot number and control point number";
    if (n < order + 1) return "fewer control points than order + 1";
    for (int64_t i = 0; i + 1 < nknots; ++i)
        if (knots[i + 1] < knots[i]) return "knot vector must be non-decreasing";
    s.periodic = periodic;
    s.uniform = false;
    s.n = n;
    s.knots.assign(knots, knots + nknots);
    set_range(order, s);
    s.dx = 0.0;
    fill_recip(order, s);
    return "";
}

}  // namespace intp_amd
