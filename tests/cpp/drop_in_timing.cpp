// Per-call cost of the C++ drop-in's scalar members, after the pattern of
// the reference's eval-proxy timing (test/src/interpolation-eval-proxy-test.cpp:
// 2^15 random coordinates sorted by cell, evaluated 32 times directly and
// 32 times through eval_proxy objects built once per coordinate).
// Meshes of about 2^18 points in 1D to 4D, even dimensions periodic,
// values a product of cosines; orders 1, 3 and 5.
//
// The scalar members run on the calling thread from a host copy of the
// control points (downloaded on the first scalar call); the proxies hold
// their precomputed weights and stencil offsets on the host.  Given the
// path of the compiled reference (oracle/_ref/libintp_ref.so, loaded with
// dlopen), the reference's own operator() is timed on the same mesh and
// points, one thread, for comparison.
// usage: drop_in_timing [libintp_ref.so]
// Prints one JSON object per configuration; needs a GPU for the builds.
#include <intp/Interpolation.hpp>

#include <dlfcn.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <utility>
#include <vector>

namespace {

constexpr std::size_t kPoints = std::size_t{1} << 15;
constexpr int kRepeat = 32;

using Clock = std::chrono::steady_clock;

// The compiled reference's C entry points (oracle/ref_driver.cpp).
struct Reference {
    void* (*build)(int, int, const int*, const int64_t*, const double*, const double*, const double* const*,
                   const double*) = nullptr;
    int (*eval)(void*, int64_t, const double*, int64_t, const int64_t*, double*, int, int) = nullptr;
    void (*release)(void*) = nullptr;
    bool ok() const { return build && eval && release; }
};
Reference g_ref;

template <std::size_t>
std::pair<double, double> full_period() {
    return {-M_PI, M_PI};
}

double ns_since(Clock::time_point t0, double calls) {
    return std::chrono::duration<double, std::nano>(Clock::now() - t0).count() / calls;
}

template <std::size_t D, std::size_t K, std::size_t... I>
void run(std::index_sequence<I...>, const std::array<std::size_t, D>& n) {
    const std::array<bool, D> periodic{(I % 2 == 0)...};
    intp::Mesh<double, D> mesh(n[I]...);
    std::array<double, D> dt{(2 * M_PI / static_cast<double>(n[I]))...};
    for (std::size_t i = 0; i < mesh.size(); ++i) {
        std::array<std::size_t, D> idx{};
        std::size_t rem = i;
        for (std::size_t d = D; d-- > 0;) {
            idx[d] = rem % n[d];
            rem /= n[d];
        }
        double v = 1.0;
        for (std::size_t d = 0; d < D; ++d) v *= std::cos(static_cast<double>(idx[d]) * dt[d] - M_PI);
        mesh(idx) = v;
    }
    std::mt19937_64 gen(12345 + D * 10 + K);
    std::uniform_real_distribution<double> uni(-M_PI, M_PI);
    std::vector<std::array<double, D>> pts(kPoints);
    for (auto& p : pts)
        for (auto& c : p) c = uni(gen);
    // sorted by cell, as the reference's test does
    std::sort(pts.begin(), pts.end(), [&](const auto& a, const auto& b) {
        for (std::size_t d = 0; d < D; ++d) {
            const auto ca = static_cast<long>((a[d] + M_PI) / dt[d]), cb = static_cast<long>((b[d] + M_PI) / dt[d]);
            if (ca != cb) return ca < cb;
        }
        return false;
    });

    const intp::InterpolationFunctionTemplate<double, D, K> tmpl(periodic, mesh.dimension(), full_period<I>()...);
    auto t0 = Clock::now();
    std::vector<typename decltype(tmpl)::eval_proxy_t> proxies;
    proxies.reserve(kPoints);
    for (const auto& x : pts) proxies.push_back(tmpl.eval_proxy(x));
    const double create_ns = ns_since(t0, kPoints);
    const auto fn = tmpl.interpolate(mesh);
    (void)fn(pts[0]);  // first scalar call: downloads the host copy

    double acc_direct = 0, acc_proxy = 0, acc_deriv = 0;
    t0 = Clock::now();
    for (int r = 0; r < kRepeat; ++r)
        for (const auto& x : pts) acc_direct += fn(x);
    const double direct_ns = ns_since(t0, static_cast<double>(kPoints) * kRepeat);
    t0 = Clock::now();
    for (int r = 0; r < kRepeat; ++r)
        for (const auto& p : proxies) acc_proxy += p(fn);
    const double proxy_ns = ns_since(t0, static_cast<double>(kPoints) * kRepeat);
    std::array<std::size_t, D> first{};
    first[0] = 1;
    t0 = Clock::now();
    for (int r = 0; r < kRepeat; ++r)
        for (const auto& x : pts) acc_deriv += fn.derivative(x, first);
    const double deriv_ns = ns_since(t0, static_cast<double>(kPoints) * kRepeat);
    // the reference's operator() on the same mesh and points (one thread)
    double ref_ns = -1, ref_max_diff = -1;
    if (g_ref.ok()) {
        std::array<int, D> per{};
        std::array<int64_t, D> nn{};
        std::array<double, D> lo{}, hi{};
        std::array<const double*, D> xc{};
        for (std::size_t d = 0; d < D; ++d) {
            per[d] = periodic[d] ? 1 : 0;
            nn[d] = static_cast<int64_t>(n[d]);
            lo[d] = -M_PI;
            hi[d] = M_PI;
        }
        std::vector<double> values(mesh.begin(), mesh.end());
        void* h = g_ref.build(static_cast<int>(D), static_cast<int>(K), per.data(), nn.data(), lo.data(), hi.data(),
                              xc.data(), values.data());
        if (h) {
            std::vector<double> xs(kPoints * D), out(kPoints), mine(kPoints);
            for (std::size_t i = 0; i < kPoints; ++i)
                for (std::size_t d = 0; d < D; ++d) xs[i * D + d] = pts[i][d];
            const std::array<int64_t, D> zero{};
            (void)g_ref.eval(h, static_cast<int64_t>(kPoints), xs.data(), 1, zero.data(), out.data(), 1, 0);
            t0 = Clock::now();
            for (int r = 0; r < kRepeat; ++r)
                (void)g_ref.eval(h, static_cast<int64_t>(kPoints), xs.data(), 1, zero.data(), out.data(), 1, 0);
            ref_ns = ns_since(t0, static_cast<double>(kPoints) * kRepeat);
            ref_max_diff = 0;
            for (std::size_t i = 0; i < kPoints; ++i)
                ref_max_diff = std::max(ref_max_diff, std::fabs(fn(pts[i]) - out[i]));
            g_ref.release(h);
        }
    }
    std::printf(
        "{\"dim\": %zu, \"order\": %zu, \"mesh\": %zu, \"points\": %zu, \"repeat\": %d, \"operator_ns\": %.1f, "
        "\"derivative_ns\": %.1f, \"proxy_apply_ns\": %.1f, \"proxy_create_ns\": %.1f, \"proxy_minus_direct\": %.3e, "
        "\"reference_operator_ns\": %.1f, \"max_abs_diff_vs_reference\": %.3e}\n",
        D, K, mesh.size(), kPoints, kRepeat, direct_ns, deriv_ns, proxy_ns, create_ns,
        (acc_proxy - acc_direct) / kRepeat, ref_ns, ref_max_diff);
    std::fflush(stdout);
    (void)acc_deriv;
}

template <std::size_t K>
void orders_for_all_dims() {
    run<1, K>(std::make_index_sequence<1>{}, {std::size_t{1} << 18});
    run<2, K>(std::make_index_sequence<2>{}, {512, 512});
    run<3, K>(std::make_index_sequence<3>{}, {64, 64, 64});
    run<4, K>(std::make_index_sequence<4>{}, {24, 24, 24, 24});
}

}  // namespace

int main(int argc, char** argv) {
    if (argc > 1) {
        if (void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL)) {
            g_ref.build = reinterpret_cast<decltype(g_ref.build)>(dlsym(lib, "ref_build"));
            g_ref.eval = reinterpret_cast<decltype(g_ref.eval)>(dlsym(lib, "ref_eval"));
            g_ref.release = reinterpret_cast<decltype(g_ref.release)>(dlsym(lib, "ref_free"));
        } else {
            std::fprintf(stderr, "cannot load %s: %s\n", argv[1], dlerror());
        }
    }
    orders_for_all_dims<1>();
    orders_for_all_dims<3>();
    orders_for_all_dims<5>();
    return 0;
}
