// ref_driver.cpp -- C-ABI shim over the UNMODIFIED reference headers
// (12ff54e/BSplineInterpolation src/include/*.hpp), compiled by
// oracle/Makefile into oracle/_ref/libintp_ref.so.  Test infrastructure
// only: it instantiates the reference templates so the oracle (and, as the
// CPU baseline, bench.py) can run the reference's own code on the same
// inputs.  Built with the reference's Release macro set
// (test/CMakeLists.txt:43-47): INTP_MULTITHREAD, INTP_PERIODIC_NO_DUMMY_POINT,
// INTP_CELL_LAYOUT.
#include <Interpolation.hpp>

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct RefBase {
    virtual ~RefBase() = default;
    virtual int dim() const = 0;
    virtual void eval(int64_t npts, const double* xs, int64_t nderiv,
                      const int64_t* orders, double* out, int nthreads,
                      int bounds, int* bad) const = 0;
    virtual void range(int d, double* lo, double* hi) const = 0;
};

template <std::size_t D, std::size_t K>
struct RefFn final : RefBase {
    using F = intp::InterpolationFunction<double, D, K>;
    F f;
    explicit RefFn(F&& g) : f(std::move(g)) {}
    int dim() const override { return static_cast<int>(D); }
    void range(int d, double* lo, double* hi) const override {
        *lo = f.range(static_cast<std::size_t>(d)).first;
        *hi = f.range(static_cast<std::size_t>(d)).second;
    }
    void eval_range(int64_t b, int64_t e, int64_t npts, const double* xs,
                    int64_t nderiv, const int64_t* orders, double* out,
                    int bounds, int* bad) const {
        for (int64_t p = b; p < e; ++p) {
            std::array<double, D> c;
            for (std::size_t d = 0; d < D; ++d) c[d] = xs[p * (int64_t)D + (int64_t)d];
            for (int64_t o = 0; o < nderiv; ++o) {
                std::array<std::size_t, D> r;
                bool zero = true;
                for (std::size_t d = 0; d < D; ++d) {
                    r[d] = static_cast<std::size_t>(orders[o * (int64_t)D + (int64_t)d]);
                    zero = zero && r[d] == 0;
                }
                double v = 0;
                try {
                    if (zero)
                        v = bounds ? f.at(c) : f(c);
                    else
                        v = bounds ? f.derivative_at(c, r) : f.derivative(c, r);
                } catch (const std::domain_error&) {
                    *bad = 1;
                }
                out[o * npts + p] = v;
            }
        }
    }
    void eval(int64_t npts, const double* xs, int64_t nderiv,
              const int64_t* orders, double* out, int nthreads, int bounds,
              int* bad) const override {
        if (nthreads <= 1) {
            eval_range(0, npts, npts, xs, nderiv, orders, out, bounds, bad);
            return;
        }
        std::vector<std::thread> th;
        std::vector<int> bads(static_cast<std::size_t>(nthreads), 0);
        const int64_t chunk = (npts + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; ++t) {
            const int64_t b = std::min(npts, t * chunk), e = std::min(npts, (t + 1) * chunk);
            th.emplace_back([=, this, &bads] {
                eval_range(b, e, npts, xs, nderiv, orders, out, bounds,
                           &bads[static_cast<std::size_t>(t)]);
            });
        }
        for (auto& x : th) x.join();
        for (int v : bads) *bad |= v;
    }
};

using Range = std::pair<double, double>;
using CoordRange = std::pair<const double*, const double*>;

template <std::size_t D>
intp::Mesh<double, D> make_mesh(const int64_t* n, const double* f) {
    std::array<std::size_t, D> dims;
    std::size_t tot = 1;
    for (std::size_t d = 0; d < D; ++d) {
        dims[d] = static_cast<std::size_t>(n[d]);
        tot *= dims[d];
    }
    intp::Mesh<double, D> m(dims);
    for (std::size_t i = 0; i < tot; ++i) {
        m(m.dimension().dimwise_indices(i)) = f[i];
    }
    return m;
}

template <std::size_t D>
std::array<bool, D> make_per(const int* periodic) {
    std::array<bool, D> p{};
    for (std::size_t d = 0; d < D; ++d) p[d] = periodic[d] != 0;
    return p;
}

template <std::size_t K>
RefBase* build_k(int dim, const int* per, const int64_t* n, const double* lo,
                 const double* hi, const double* const* xc, const double* f) {
    auto cr = [&](int d) {
        return CoordRange{xc[d], xc[d] + n[d] + (per[d] ? 1 : 0)};
    };
    auto ur = [&](int d) { return Range{lo[d], hi[d]}; };
    const bool nu0 = xc && xc[0];
    const bool nu1 = xc && dim > 1 && xc[1];
    const bool nu2 = xc && dim > 2 && xc[2];
    if (dim == 1) {
        auto m = make_mesh<1>(n, f);
        if (nu0)
            return new RefFn<1, K>(intp::InterpolationFunction<double, 1, K>(make_per<1>(per), m, cr(0)));
        return new RefFn<1, K>(intp::InterpolationFunction<double, 1, K>(make_per<1>(per), m, ur(0)));
    }
    if (dim == 2) {
        auto m = make_mesh<2>(n, f);
        using F2 = intp::InterpolationFunction<double, 2, K>;
        if (!nu0 && !nu1) return new RefFn<2, K>(F2(make_per<2>(per), m, ur(0), ur(1)));
        if (!nu0 && nu1) return new RefFn<2, K>(F2(make_per<2>(per), m, ur(0), cr(1)));
        if (nu0 && !nu1) return new RefFn<2, K>(F2(make_per<2>(per), m, cr(0), ur(1)));
        return new RefFn<2, K>(F2(make_per<2>(per), m, cr(0), cr(1)));
    }
    if (dim == 3) {
        auto m = make_mesh<3>(n, f);
        using F3 = intp::InterpolationFunction<double, 3, K>;
        if (!nu0 && !nu1 && !nu2) return new RefFn<3, K>(F3(make_per<3>(per), m, ur(0), ur(1), ur(2)));
        if (nu0 && nu1 && nu2) return new RefFn<3, K>(F3(make_per<3>(per), m, cr(0), cr(1), cr(2)));
        throw std::invalid_argument("mixed uniform/non-uniform 3D not instantiated");
    }
    if (dim == 4) {
        // uniform ranges only (the 4D pinning cases)
        if (xc && (xc[0] || xc[1] || xc[2] || xc[3]))
            throw std::invalid_argument("non-uniform 4D not instantiated");
        auto m = make_mesh<4>(n, f);
        using F4 = intp::InterpolationFunction<double, 4, K>;
        return new RefFn<4, K>(F4(make_per<4>(per), m, ur(0), ur(1), ur(2), ur(3)));
    }
    throw std::invalid_argument("dim not instantiated");
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Builds InterpolationFunction<double, dim, order>(periodicity, mesh, ranges...)
// (Interpolation.hpp:92-101).  xc[d] != NULL selects the non-uniform
// (iterator pair) range for that dim.
void* ref_build(int dim, int order, const int* per, const int64_t* n,
                const double* lo, const double* hi, const double* const* xc,
                const double* f) {
    try {
        switch (order) {
            case 0: return build_k<0>(dim, per, n, lo, hi, xc, f);
            case 1: return build_k<1>(dim, per, n, lo, hi, xc, f);
            case 2: return build_k<2>(dim, per, n, lo, hi, xc, f);
            case 3: return build_k<3>(dim, per, n, lo, hi, xc, f);
            case 4: return build_k<4>(dim, per, n, lo, hi, xc, f);
            case 5: return build_k<5>(dim, per, n, lo, hi, xc, f);
            case 6: return build_k<6>(dim, per, n, lo, hi, xc, f);
            case 7: return build_k<7>(dim, per, n, lo, hi, xc, f);
            default: throw std::invalid_argument("order not instantiated");
        }
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// operator() / derivative() (bounds=0) or at() / derivative_at() (bounds=1)
// per point; returns -1 if any point raised std::domain_error.
int ref_eval(void* h, int64_t npts, const double* xs, int64_t nderiv,
             const int64_t* orders, double* out, int nthreads, int bounds) {
    int bad = 0;
    static_cast<RefBase*>(h)->eval(npts, xs, nderiv, orders, out, nthreads, bounds, &bad);
    return bad ? -1 : 0;
}

void ref_range(void* h, int d, double* lo, double* hi) {
    static_cast<RefBase*>(h)->range(d, lo, hi);
}

void ref_free(void* h) { delete static_cast<RefBase*>(h); }

// Explicit-knot BSpline<double, D, 3> (BSpline.hpp:188-210), value or
// derivative_at (BSpline.hpp:393-550) per point, for the bspline-test
// known-answer vectors.
int ref_bspline3_eval(int dim, const int* per, const int64_t* n,
                      const double* const* knots, const int64_t* nknots,
                      const double* ctrl, int64_t npts, const double* xs,
                      const int64_t* order, double* out) {
    try {
        if (dim == 1) {
            using S = intp::BSpline<double, 1, 3>;
            typename S::ControlPointContainer cp(std::make_pair(ctrl, ctrl + n[0]));
            S s(make_per<1>(per), cp, std::make_pair(knots[0], knots[0] + nknots[0]));
            for (int64_t p = 0; p < npts; ++p) {
                std::array<std::pair<double, std::size_t>, 1> c{{{xs[p], (std::size_t)order[0]}}};
                out[p] = s.derivative_at(c);
            }
            return 0;
        }
        if (dim == 2) {
            using S = intp::BSpline<double, 2, 3>;
            typename S::ControlPointContainer cp(static_cast<std::size_t>(n[0]), static_cast<std::size_t>(n[1]));
            for (int64_t i = 0; i < n[0]; ++i)
                for (int64_t j = 0; j < n[1]; ++j) cp(i, j) = ctrl[i * n[1] + j];
            S s(make_per<2>(per), cp, std::make_pair(knots[0], knots[0] + nknots[0]),
                std::make_pair(knots[1], knots[1] + nknots[1]));
            for (int64_t p = 0; p < npts; ++p) {
                std::array<std::pair<double, std::size_t>, 2> c{
                    {{xs[2 * p], (std::size_t)order[0]}, {xs[2 * p + 1], (std::size_t)order[1]}}};
                out[p] = s.derivative_at(c);
            }
            return 0;
        }
        return -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
