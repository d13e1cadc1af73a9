// C++ drop-in test: the reference's interpolation known answers (generated
// blocks, mma_cases.inc) plus the behaviour of the class surface that the
// reference's tests rely on -- std::domain_error from at(), templates
// reused over several meshes, eval_proxy, value semantics of copies.
// Runs on a GPU through libintp_amd.so; prints one line per check and
// exits non-zero on any failure.
#include <intp/Interpolation.hpp>

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
    bool same = true;
    for (int i = 0; i < 1000; i += 37) {
        const std::array<double, 2> x{pts[2 * i], pts[2 * i + 1]};
        same = same && out[i] == fn(x) && out[1000 + i] == fn.derivative(x, 1, 0) &&
               out[2000 + i] == fn.derivative(x, 0, 1);
    }
    expect("batched outputs == scalar calls", same);
}

}  // namespace

int main() {
#include "mma_cases.inc"
    surface_checks();
    std::printf("%d of %d checks failed\n", failures, checks);
    return failures == 0 ? 0 : 1;
}
