// setup.hpp -- host-side preparation of one grid dimension: knot vector,
// range, collocation matrix and its no-pivot (cyclic) band LU, converted to
// the row-oriented factor arrays the GPU line solver streams through.
// Everything here is O(n * order^2) per dimension, done once per template.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace intp_amd {

// Row-oriented LU factors of one dimension's collocation matrix A = L U,
// L unit lower with p sub-diagonals, U upper with q super-diagonals, plus
// the cyclic fill of a periodic matrix:
//   lower[i*p + m-1] = L(i, i-m)          (m = 1..p)
//   upper[i*q + m-1] = U(i, i+m)          (m = 1..q)
//   diag[i]          = U(i, i)
//   bottom[r*n + j]  = L(n-q+r, j), j < n-q+r-p   (periodic only)
//   right[c*n + i]   = U(i, n-p+c), i < n-p+c-q   (periodic only)
struct LineFactors {
    int64_t n = 0;
    int p = 0, q = 0;
    bool cyclic = false;
    std::vector<double> lower, upper, diag, bottom, right;
};

struct DimSetup {
    bool periodic = false;
    bool uniform = true;
    int64_t n = 0;                // stored values = control points
    std::vector<double> knots;    // n + order + 1 (+ order, periodic)
    std::vector<double> nodes;    // non-uniform node coordinates
    double lo = 0, hi = 0, dx = 0;  // range and uniform spacing
    LineFactors lu;
    // The collocation matrix itself (before LU), for residuals of the
    // refined-FP32 build: colloc[i*(2w+1) + (m+w)] = A(i, i+m), columns
    // taken mod n on periodic dimensions (zero outside a clamped matrix),
    // m = -w..w, w = colloc_w.
    int colloc_w = 0;
    std::vector<double> colloc;
    // The LU factors rounded to FP32 (same layout as `lu`).
    std::vector<float> lower32, upper32, diag32, bottom32, right32;
    // reciprocal knot differences for the basis recursion, per span mu:
    // recip[mu*T + i*(i-1)/2 + a-1] = 1 / (t[mu+a] - t[mu+a-i]),
    // i = 1..order, a = 1..i, T = order*(order+1)/2
    std::vector<double> recip;

    // Uniform dimensions only: the knot formula and the basis polynomials.
    // t(i) = xmin for i <= order (clamped), xmax for i >= nt-order-1
    // (clamped), else xmin + (i - half) * dx.
    double xmin = 0, xmax = 0, half = 0;
    // Spans whose 2*order knots all follow the formula share one set of
    // basis polynomials (class 0).  The first `nleft` spans (mu = order..)
    // and the last `nright` spans of a clamped dimension get their own
    // classes: 1 + (mu - order) and 1 + nleft + (last - mu).
    int nleft = 0, nright = 0, nclass = 1;
    // poly[(c*(order+1) + j)*(order+1) + p]: coefficient of u^p of the j-th
    // non-zero B-spline on a span of class c, u = (x - t(mu)) / dx.
    std::vector<double> poly;
};

// Returns an empty string on success, else the reason.
std::string setup_dimension(int order, bool periodic, int64_t n, double xmin, double xmax,
                            const double* nodes, DimSetup& out);
// For explicit knot vectors (no collocation matrix).
std::string setup_from_knots(int order, bool periodic, int64_t n, const double* knots,
                             int64_t nknots, DimSetup& out);
// Cox-de Boor values of the order+1 B-splines non-zero on span mu at x;
// for spline_order < order the lower-order values sit right-aligned.
void basis_values(const double* t, int64_t mu, double x, int order, int spline_order, double* out);
int64_t knot_count(int order, bool periodic, int64_t n);

}  // namespace intp_amd
