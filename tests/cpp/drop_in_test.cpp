// C++ drop-in test: the reference's interpolation and explicit-knot
// BSpline known answers (generated blocks, mma_cases.inc) plus the behaviour of the class surface that the
// reference's tests rely on -- std::domain_error from at(), templates
// reused over several meshes, eval_proxy, value semantics of copies.
// Runs on a GPU through libintp_amd.so; prints one line per check and
// exits non-zero on any failure.
#include <intp/Interpolation.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <utility>
#include <vector>

namespace {

int failures = 0;
int checks = 0;

// norm-wise relative L2 error (the reference's rel_err, test/src/include/rel_err.hpp:11-49)
double rel_err(const double* got, const double* want, std::size_t n) {
    double num = 0, den = 0;
    for (std::size_t i = 0; i < n; ++i) {
        num += (got[i] - want[i]) * (got[i] - want[i]);
        den += want[i] * want[i];
    }
    return std::sqrt(num / den);
}

void report(const char* name, const double* got, const double* want, std::size_t n, double tol) {
    const double e = rel_err(got, want, n);
    const bool ok = e < tol;
    ++checks;
    if (!ok) ++failures;
    std::printf("%-40s rel_err %.3e  (tol %.0e)  %s\n", name, e, tol, ok ? "ok" : "FAILED");
}

void expect(const char* name, bool ok) {
    ++checks;
    if (!ok) ++failures;
    std::printf("%-40s %s\n", name, ok ? "ok" : "FAILED");
}

void surface_checks() {
    // a smooth 2D function sampled on 24 x 20, y periodic
    intp::Mesh<double, 2> f(24, 20);
    intp::Mesh<double, 2> g(24, 20);
    for (std::size_t i = 0; i < 24; ++i)
        for (std::size_t j = 0; j < 20; ++j) {
            const double x = i / 23.0, y = 2.0 * M_PI * j / 20.0;
            f(i, j) = std::exp(-x) * std::sin(y);
            g(i, j) = x * x + std::cos(2.0 * y);
        }
    const auto xr = std::make_pair(0.0, 1.0);
    const auto yr = std::make_pair(0.0, 2.0 * M_PI);
    const intp::InterpolationFunction<double, 2, 3> fn({false, true}, f, xr, yr);

    // properties (Interpolation.hpp:248-267)
    expect("periodicity/uniform/range/get_order",
           !fn.periodicity(0) && fn.periodicity(1) && fn.uniform(0) && fn.uniform(1) && fn.range(0).first == 0.0 &&
               fn.range(0).second == 1.0 && fn.range(1).second == 2.0 * M_PI && fn.get_order() == 3);

    // at() / derivative_at() out of range: std::domain_error; periodic
    // dimensions are never out of range (Interpolation.hpp:482-491)
    bool thrown = false;
    try {
        (void)fn.at(1.5, 0.3);
    } catch (const std::domain_error&) {
        thrown = true;
    }
    expect("at() outside a clamped range throws", thrown);
    thrown = false;
    try {
        (void)fn.derivative_at({-0.25, 0.3}, 1, 0);
    } catch (const std::domain_error&) {
        thrown = true;
    }
    expect("derivative_at() outside throws", thrown);
    bool fine = true;
    try {
        (void)fn.at(0.5, 9.0);
    } catch (const std::exception&) {
        fine = false;
    }
    expect("at() wraps a periodic coordinate", fine);
    // without the check the boundary polynomial extrapolates
    expect("operator() extrapolates", std::isfinite(fn(1.2, 0.4)));

    // the three derivative argument forms agree
    const std::array<double, 2> c{0.37, 1.9};
    const double d1 = fn.derivative(c, std::array<std::size_t, 2>{1, 2});
    const double d2 = fn.derivative(c, 1, 2);
    const double d3 = fn.derivative(std::make_pair(c[0], 1), std::make_pair(c[1], 2));
    expect("derivative forms agree", d1 == d2 && d2 == d3);

    // template: factor once, interpolate many meshes (InterpolationTemplate.hpp:118-143)
    const intp::InterpolationFunctionTemplate<double, 2, 3> tmpl({false, true}, intp::MeshDimension<2>(24, 20), xr, yr);
    const auto fa = tmpl.interpolate(f);
    const auto ga = tmpl.interpolate(g);
    const intp::InterpolationFunction<double, 2, 3> gn({false, true}, g, xr, yr);
    expect("template == constructor (control points)",
           fa.control_points() == fn.control_points() && ga.control_points() == gn.control_points());
    intp::InterpolationFunction<double, 2, 3> filled;
    tmpl.interpolate(filled, g);
    expect("template fills an empty function", filled(0.2, 0.7) == gn(0.2, 0.7));

    // eval_proxy: fixed coordinate, any function on the grid (Interpolation.hpp:493-506)
    const auto proxy = tmpl.eval_proxy(0.61, 4.2);
    expect("eval_proxy", proxy(fa) == fn(0.61, 4.2) && proxy(ga) == gn(0.61, 4.2));
    const auto proxy2 = fn.eval_proxy({0.1, 0.2});
    expect("function eval_proxy", proxy2(gn) == gn(0.1, 0.2));

    // value semantics: a copy outlives the original
    intp::InterpolationFunction<double, 2, 3> copy;
    double before = 0;
    {
        const intp::InterpolationFunction<double, 2, 3> tmp({false, true}, f, xr, yr);
        before = tmp(0.3, 0.3);
        copy = tmp;
    }
    expect("copy outlives original", copy(0.3, 0.3) == before);

    // an empty function refuses to evaluate
    thrown = false;
    try {
        const intp::InterpolationFunction<double, 2, 3> empty;
        (void)empty(0.1, 0.1);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    expect("empty function throws", thrown);

    // 1D template over iterator ranges (InterpolationTemplate.hpp:583-604)
    std::vector<double> v1{1, 3, 2, 5, 4, 6, 8, 7};
    const intp::InterpolationFunctionTemplate1D<3> t1(v1.size());
    const auto f1 = t1.interpolate(intp::util::get_range(v1));
    const intp::InterpolationFunction1D<3> f1d(intp::util::get_range(v1));
    expect("1D template == 1D function", f1(2.5) == f1d(2.5) && std::fabs(f1(3.0) - 5.0) < 1e-14);

    // batched evaluation with several outputs: value and both partials
    std::vector<double> pts, out(3 * 1000);
    for (int i = 0; i < 1000; ++i) {
        pts.push_back(i / 999.0);
        pts.push_back(6.0 * i / 999.0);
    }
    const std::size_t orders[6] = {0, 0, 1, 0, 0, 1};
    fn.eval(1000, pts.data(), out.data(), 3, orders);
    // batched calls run on the GPU (per-class basis polynomials), scalar
    // calls on the host copy (Cox-de Boor): equal to round-off
    double worst = 0;
    for (int i = 0; i < 1000; i += 37) {
        const std::array<double, 2> x{pts[2 * i], pts[2 * i + 1]};
        const double s[3] = {fn(x), fn.derivative(x, 1, 0), fn.derivative(x, 0, 1)};
        for (int o = 0; o < 3; ++o)
            worst = std::max(worst, std::fabs(out[o * 1000 + i] - s[o]) / std::max(1.0, std::fabs(s[o])));
    }
    std::printf("batched vs scalar: max relative difference %.3e\n", worst);
    expect("batched outputs == scalar calls (1e-13)", worst < 1e-13);

    // spline(): read-only view of the BSpline (Interpolation.hpp:264)
    const auto& sp = fn.spline();
    const std::vector<double> k0 = fn.knots(0);
    expect("spline() knots", sp.knots_num(0) == k0.size() && std::equal(sp.knots_begin(0), sp.knots_end(0), k0.begin()) &&
                                 *(sp.knots_end(1) - 1) == fn.knots(1).back());
    const std::vector<double> cp = fn.control_points();
    expect("spline() control points", sp.control_points().size() == cp.size() &&
                                          std::equal(sp.control_points().begin(), sp.control_points().end(), cp.begin()) &&
                                          sp.control_points().dim_size(0) == 24);
    expect("spline() properties", sp.periodicity(1) && !sp.periodicity(0) && sp.get_order() == 3 &&
                                      sp.range(1).second == fn.range(1).second);
    expect("spline() evaluation", sp(0.3, 2.0) == fn(0.3, 2.0) &&
                                      sp.derivative_at({std::make_pair(0.3, std::size_t{1}), std::make_pair(2.0, std::size_t{2})}) ==
                                          fn.derivative({0.3, 2.0}, 1, 2));
}

// A fixed-size vector value type, as the reference's Vec<D, T>
// (test/src/include/Vec.hpp): static dim, value_type, operator[].
struct Vec2f {
    static constexpr std::size_t dim = 2;
    using value_type = float;
    std::array<float, 2> c{};
    Vec2f() = default;
    Vec2f(double x, double y) : c{static_cast<float>(x), static_cast<float>(y)} {}
    float& operator[](std::size_t i) { return c[i]; }
    const float& operator[](std::size_t i) const { return c[i]; }
    float mag() const { return std::sqrt(c[0] * c[0] + c[1] * c[1]); }
};

void value_type_checks() {
    // float values and coordinates: the refined-FP32 form
    intp::Mesh<float, 2> f(30, 26);
    intp::Mesh<double, 2> g(30, 26);
    for (std::size_t i = 0; i < 30; ++i)
        for (std::size_t j = 0; j < 26; ++j) {
            const double x = i / 29.0, y = j / 25.0;
            g(i, j) = std::sin(3.0 * x) * std::cos(2.0 * y) + x * y;
            f(i, j) = static_cast<float>(g(i, j));
        }
    const intp::InterpolationFunction<float, 2, 3, float> ff(f, std::make_pair(0.0f, 1.0f), std::make_pair(0.0f, 1.0f));
    const intp::InterpolationFunction<double, 2, 3> fd(g, std::make_pair(0.0, 1.0), std::make_pair(0.0, 1.0));
    double worst = 0;
    for (int i = 0; i < 200; ++i) {
        const float x = static_cast<float>((i * 37 % 200) / 199.0), y = static_cast<float>(i / 199.0);
        worst = std::max(worst, std::fabs(static_cast<double>(ff(x, y)) - fd(static_cast<double>(x), static_cast<double>(y))));
        worst = std::max(worst, std::fabs(static_cast<double>(ff.derivative({x, y}, 1, 0)) -
                                          fd.derivative({static_cast<double>(x), static_cast<double>(y)}, 1, 0)) / 10.0);
    }
    std::printf("float function vs double function: max difference %.3e\n", worst);
    expect("InterpolationFunction<float, 2, 3, float> (1e-5)", worst < 1e-5);

    // 2D points on a circle, periodic, vector-valued
    // (interpolation-test.cpp:674-703, tolerance 1e-5)
    // (the closing point repeats the first under the default layout)
    std::vector<Vec2f> pts;
    const std::size_t pts_n = 31;
#ifdef INTP_PERIODIC_NO_DUMMY_POINT
    for (std::size_t i = 0; i < pts_n; ++i) {
#else
    for (std::size_t i = 0; i <= pts_n; ++i) {
#endif
        const double theta = 2. * M_PI * static_cast<double>(i) / pts_n;
        pts.emplace_back(std::cos(theta), std::sin(theta));
    }
    const intp::InterpolationFunction1D<3, Vec2f, float> circle(intp::util::get_range(pts), true);
    const std::size_t sample_n = 1024;
    float err = 0;
    for (std::size_t i = 0; i < sample_n; ++i) {
        const float theta = 2.f * static_cast<float>(M_PI) * static_cast<float>(i) / sample_n;
        err += std::abs(circle(theta).mag() - 1.f);
    }
    err /= sample_n;
    std::printf("circle of Vec2f: mean |mag - 1| = %.3e\n", err);
    expect("InterpolationFunction1D<3, Vec2f, float> circle (1e-5)", err < 1.e-5f);
    // the vector spline's components are the scalar splines of the components
    std::vector<float> xs;
    for (const auto& p : pts) xs.push_back(p[0]);
    const intp::InterpolationFunction1D<3, float, float> cx(intp::util::get_range(xs), true);
    expect("vector components == scalar splines", circle(1.3f)[0] == cx(1.3f));
    const auto proxy = intp::InterpolationFunctionTemplate1D<3, Vec2f, float>(pts_n, true).eval_proxy(2.5f);
    (void)proxy;
}

// 4D: InterpolationFunction<double, 4, 3> of a separable product against
// the products of the 1D splines of its factors (tensor-product
// interpolation of a rank-one mesh is rank one), value and one partial.
void high_dim_checks() {
    const std::array<std::size_t, 4> n{8, 7, 6, 9};
    const std::array<bool, 4> per{true, false, false, true};
    std::array<std::vector<double>, 4> g;
    for (std::size_t d = 0; d < 4; ++d)
        for (std::size_t i = 0; i < n[d]; ++i) g[d].push_back(std::cos(0.7 * static_cast<double>(i) + d) + 0.1 * d);
    intp::Mesh<double, 4> m(n[0], n[1], n[2], n[3]);
    for (std::size_t a = 0; a < n[0]; ++a)
        for (std::size_t b = 0; b < n[1]; ++b)
            for (std::size_t c = 0; c < n[2]; ++c)
                for (std::size_t e = 0; e < n[3]; ++e) m(a, b, c, e) = g[0][a] * g[1][b] * g[2][c] * g[3][e];
    const auto r = std::make_pair(0.0, 1.0);
    const intp::InterpolationFunction<double, 4, 3> f4(per, m, r, r, r, r);
    const std::array<intp::InterpolationFunction1D<3>, 4> f1{
        intp::InterpolationFunction1D<3>(r, intp::util::get_range(g[0]), per[0]),
        intp::InterpolationFunction1D<3>(r, intp::util::get_range(g[1]), per[1]),
        intp::InterpolationFunction1D<3>(r, intp::util::get_range(g[2]), per[2]),
        intp::InterpolationFunction1D<3>(r, intp::util::get_range(g[3]), per[3])};
    std::vector<double> got, want, got_d, want_d;
    for (int i = 0; i < 100; ++i) {
        const std::array<double, 4> x{(i * 37 % 100) / 99.0, (i * 53 % 100) / 99.0, (i * 71 % 100) / 99.0, i / 99.0};
        got.push_back(f4(x));
        want.push_back(f1[0](x[0]) * f1[1](x[1]) * f1[2](x[2]) * f1[3](x[3]));
        got_d.push_back(f4.derivative(x, std::array<std::size_t, 4>{0, 1, 0, 0}));
        want_d.push_back(f1[0](x[0]) * f1[1].derivative(std::make_pair(x[1], std::size_t{1})) * f1[2](x[2]) *
                         f1[3](x[3]));
    }
    report("4D value = product of 1D splines", got.data(), want.data(), got.size(), 1e-12);
    report("4D d/dx1 = product with 1D derivative", got_d.data(), want_d.data(), got_d.size(), 1e-12);
}

// The explicit-knot BSpline constructors in the reference test's own
// spelling (bspline-test.cpp:49-58): a 1D control mesh from an iterator
// pair, the knots as an initializer list; a knot count that does not
// match the control points throws.
void bspline_checks() {
    using Spline = intp::BSpline<double, 1, 3>;
    const auto knots = {0., 0., 0., 0., .5, 1., 1., 1., 1.};
    const auto cp = {1., 16. / 3, -16. / 3, 4. / 3, 3.};
    const Spline sp(Spline::ControlPointContainer(std::make_pair(cp.begin(), cp.end())),
                    std::make_pair(knots.begin(), knots.end()));
    // the segment's end points interpolate the end control points
    expect("BSpline knots: value at 0 = first ctrl", std::abs(sp({0.0}) - 1.0) < 1e-15);
    expect("BSpline knots: value at 1 = last ctrl", std::abs(sp({1.0}) - 3.0) < 1e-15);
    expect("BSpline knots: range", sp.range(0).first == 0.0 && sp.range(0).second == 1.0);
    expect("BSpline knots: knots round trip",
           std::equal(knots.begin(), knots.end(), sp.knots_begin(0), sp.knots_end(0)));
    expect("BSpline knots: control points round trip",
           std::equal(cp.begin(), cp.end(), sp.control_points().begin()));
    bool threw = false;
    try {
        const auto short_knots = {0., 0., 0., 0., 1., 1., 1., 1.};
        const Spline bad(Spline::ControlPointContainer(std::make_pair(cp.begin(), cp.end())),
                         std::make_pair(short_knots.begin(), short_knots.end()));
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect("BSpline knots: count mismatch throws", threw);

    // the loader form (BSpline.hpp:178-186, 217-242): periodicity first,
    // then knots and control points in either order
    Spline loaded(std::array<bool, 1>{false});
    loaded.load_knots(0, Spline::KnotContainer(knots.begin(), knots.end()));
    expect("BSpline loaders: range after load_knots", loaded.range(0).first == 0.0 && loaded.range(0).second == 1.0);
    loaded.load_ctrlPts(Spline::ControlPointContainer(std::make_pair(cp.begin(), cp.end())));
    bool same = true;
    for (int i = 0; i <= 20; ++i) {
        const double x = i / 20.0;
        same = same && loaded({x}) == sp({x}) &&
               loaded.derivative_at({std::make_pair(x, std::size_t{1})}) ==
                   sp.derivative_at({std::make_pair(x, std::size_t{1})});
    }
    expect("BSpline loaders == knot constructor", same);
    // a later load rebuilds the table
    const auto cp2 = {2., 16. / 3, -16. / 3, 4. / 3, 5.};
    loaded.load_ctrlPts(Spline::ControlPointContainer(std::make_pair(cp2.begin(), cp2.end())));
    expect("BSpline loaders: reload", std::abs(loaded({0.0}) - 2.0) < 1e-15 && std::abs(loaded({1.0}) - 5.0) < 1e-15);
}

}  // namespace

int main() {
#include "mma_cases.inc"
    surface_checks();
    value_type_checks();
    high_dim_checks();
    bspline_checks();
    std::printf("%d of %d checks failed\n", failures, checks);
    return failures == 0 ? 0 : 1;
}
