// intp/Interpolation.hpp -- C++ drop-in for the reference's public surface
// (12ff54e/BSplineInterpolation, src/include/Interpolation.hpp,
// InterpolationTemplate.hpp, Mesh.hpp) on top of the C ABI of
// libintp_amd.so (include/intp_amd.h).  Building and evaluating run on the
// GPU; this header holds no numerics of its own.
//
// Kept from the reference:
//  * intp::Mesh<T, D> / intp::MeshDimension<D>: C-order value meshes
//    (Mesh.hpp:11-383);
//  * intp::InterpolationFunction<T, D, O, U> constructors (periodicity
//    array, Mesh, one (min, max) or (begin, end) coordinate pair per
//    dimension; the 1D iterator-range forms; default construction),
//    Interpolation.hpp:49-125;
//  * operator(), at, derivative and derivative_at in all three argument
//    forms, Interpolation.hpp:132-244; periodicity / uniform / range /
//    get_order, :248-267; eval_proxy, :493-506;
//  * intp::InterpolationFunctionTemplate<T, D, O, U> with interpolate() and
//    eval_proxy (InterpolationTemplate.hpp:32-176), the 1D classes
//    InterpolationFunction1D (Interpolation.hpp:509-540) and
//    InterpolationFunctionTemplate1D (InterpolationTemplate.hpp:583-604);
//  * intp::util::get_range;
//  * value semantics: a function is a cheap handle to an immutable device
//    table (copies share it, it is released with the last copy);
//  * exceptions: std::domain_error for at()/derivative_at() out of range,
//    std::bad_alloc when device memory runs out, std::runtime_error for
//    invalid arguments and HIP errors.
// Conventions: INTP_PERIODIC_NO_DUMMY_POINT (a periodic dimension stores n
// values, spacing (max - min) / n) and INTP_CELL_LAYOUT semantics.
// T = U = double (the FP64 path).
//
// Added: batched evaluation of many points in one call (eval(), host or
// device buffers), which is how the GPU is meant to be used; the scalar
// members each make one synchronous device round trip.
//
// Device selection: intp::set_device(d) before the first construction
// (default 0, or the INTP_DEVICE environment variable).
#ifndef INTP_AMD_INTERPOLATION_HPP
#define INTP_AMD_INTERPOLATION_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <iterator>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../intp_amd.h"

namespace intp {

// ---- errors and contexts -------------------------------------------------

namespace detail {

inline void check(int rc) {
    if (rc == INTP_OK) return;
    const std::string msg = intp_last_error();
    switch (rc) {
        case INTP_E_DOMAIN: throw std::domain_error(msg);
        case INTP_E_NOMEM: throw std::bad_alloc();
        default: throw std::runtime_error(std::string(intp_status_string(rc)) + ": " + msg);
    }
}

inline int& selected_device() {
    static int device = [] {
        const char* env = std::getenv("INTP_DEVICE");
        return env ? std::atoi(env) : 0;
    }();
    return device;
}

// One context per device, created on first use and kept for the process.
inline intp_ctx context() {
    static std::mutex mu;
    static std::vector<intp_ctx> contexts;
    std::lock_guard<std::mutex> lock(mu);
    const int d = selected_device();
    if (d < 0) throw std::runtime_error("negative device index");
    if (static_cast<std::size_t>(d) >= contexts.size()) contexts.resize(static_cast<std::size_t>(d) + 1, nullptr);
    if (!contexts[static_cast<std::size_t>(d)]) check(intp_ctx_create(d, &contexts[static_cast<std::size_t>(d)]));
    return contexts[static_cast<std::size_t>(d)];
}

using SplineHandle = std::shared_ptr<intp_spline_s>;
using TemplateHandle = std::shared_ptr<intp_template_s>;

inline SplineHandle own(intp_spline s) {
    return SplineHandle(s, [](intp_spline p) { intp_spline_release(p); });
}

template <typename T>
struct is_iterator {
    template <typename X>
    static auto test(int) -> decltype(typename std::iterator_traits<X>::iterator_category{}, std::true_type{});
    template <typename>
    static std::false_type test(...);
    static constexpr bool value = decltype(test<T>(0))::value && !std::is_arithmetic<T>::value;
};

// One dimension of a grid description from a (min, max) or (begin, end)
// pair (Interpolation.hpp:322-373).
template <typename A, typename B>
void describe(intp_grid_desc& desc, int d, const std::pair<A, B>& range, std::vector<std::vector<double>>& keep,
              std::true_type /* arithmetic */) {
    desc.xmin[d] = static_cast<double>(range.first);
    desc.xmax[d] = static_cast<double>(range.second);
    desc.coords[d] = nullptr;
    (void)keep;
}

template <typename A, typename B>
void describe(intp_grid_desc& desc, int d, const std::pair<A, B>& range, std::vector<std::vector<double>>& keep,
              std::false_type /* iterators */) {
    std::vector<double> nodes;
    for (auto it = range.first; it != range.second; ++it) nodes.push_back(static_cast<double>(*it));
    const std::int64_t want = desc.n[d] + (desc.periodic[d] ? 1 : 0);
    if (static_cast<std::int64_t>(nodes.size()) != want)
        throw std::runtime_error("Inconsistency between knot number and interpolated value number at dimension " +
                                 std::to_string(d));
    keep[static_cast<std::size_t>(d)] = std::move(nodes);
    desc.coords[d] = keep[static_cast<std::size_t>(d)].data();
}

template <std::size_t D, typename... Ts, std::size_t... I>
void describe_all(intp_grid_desc& desc, std::vector<std::vector<double>>& keep, std::index_sequence<I...>,
                  const std::pair<Ts, Ts>&... ranges) {
    int unused[] = {0, (describe(desc, static_cast<int>(I), ranges, keep,
                                 std::integral_constant<bool, std::is_arithmetic<Ts>::value>{}),
                        0)...};
    (void)unused;
}

}  // namespace detail

// Device later constructions use (InterpolationFunction, templates).
inline void set_device(int device) { detail::selected_device() = device; }

namespace util {

// (begin, end) of a container, the reference's util::get_range.
template <typename C>
auto get_range(C& c) -> std::pair<decltype(std::begin(c)), decltype(std::end(c))> {
    return std::make_pair(std::begin(c), std::end(c));
}

}  // namespace util

// ---- Mesh ------------------------------------------------------------------

// Sizes of a D-dimensional C-order mesh (Mesh.hpp:11-120).
template <std::size_t D>
class MeshDimension {
   public:
    using size_type = std::size_t;
    using index_type = std::array<size_type, D>;
    constexpr static size_type dim = D;

    MeshDimension() { dim_size_.fill(0); }
    MeshDimension(index_type sizes) : dim_size_(sizes) {}
    template <typename... Args, typename = typename std::enable_if<(sizeof...(Args) == D && D > 1)>::type>
    MeshDimension(Args... sizes) : dim_size_{static_cast<size_type>(sizes)...} {}
    MeshDimension(size_type n) { dim_size_.fill(n); }

    size_type size() const {
        size_type s = 1;
        for (size_type d : dim_size_) s *= d;
        return s;
    }
    size_type dim_size(size_type d) const { return dim_size_[d]; }
    size_type& dim_size(size_type d) { return dim_size_[d]; }
    // flat C-order index (last index fastest), Mesh.hpp:87-95
    size_type indexing(const index_type& ind) const {
        size_type flat = 0;
        for (size_type d = 0; d < D; ++d) flat = flat * dim_size_[d] + ind[d];
        return flat;
    }
    index_type dimwise_indices(size_type flat) const {
        index_type ind{};
        for (size_type d = D; d-- > 0;) {
            ind[d] = flat % dim_size_[d];
            flat /= dim_size_[d];
        }
        return ind;
    }
    void resize(index_type sizes) { dim_size_ = sizes; }
    operator index_type() const { return dim_size_; }

   private:
    index_type dim_size_;
};

// C-order value mesh on the host (Mesh.hpp:126-383).
template <typename T, std::size_t D>
class Mesh {
   public:
    using size_type = std::size_t;
    using val_type = T;
    using index_type = typename MeshDimension<D>::index_type;
    using iterator = typename std::vector<T>::iterator;
    using const_iterator = typename std::vector<T>::const_iterator;
    constexpr static size_type dim = D;

    explicit Mesh(MeshDimension<D> dimension) : dimension_(dimension), storage_(dimension.size()) {}
    explicit Mesh(index_type sizes) : Mesh(MeshDimension<D>(sizes)) {}
    template <typename... Args, typename = typename std::enable_if<(sizeof...(Args) == D && D > 1)>::type>
    explicit Mesh(Args... sizes) : Mesh(MeshDimension<D>(index_type{static_cast<size_type>(sizes)...})) {}
    explicit Mesh(size_type n) : Mesh(MeshDimension<D>(n)) {}
    // 1D mesh from an iterator range
    template <typename It, typename = typename std::enable_if<D == 1 && detail::is_iterator<It>::value>::type>
    explicit Mesh(std::pair<It, It> range) : dimension_(index_type{0}) {
        for (It it = range.first; it != range.second; ++it) storage_.push_back(static_cast<T>(*it));
        dimension_.resize(index_type{storage_.size()});
    }

    size_type size() const { return storage_.size(); }
    size_type dim_size(size_type d) const { return dimension_.dim_size(d); }
    const MeshDimension<D>& dimension() const { return dimension_; }

    template <typename... Indices>
    T& operator()(Indices... ind) {
        return storage_[dimension_.indexing(index_type{static_cast<size_type>(ind)...})];
    }
    template <typename... Indices>
    const T& operator()(Indices... ind) const {
        return storage_[dimension_.indexing(index_type{static_cast<size_type>(ind)...})];
    }
    T& operator()(index_type ind) { return storage_[dimension_.indexing(ind)]; }
    const T& operator()(index_type ind) const { return storage_[dimension_.indexing(ind)]; }

    T* data() { return storage_.data(); }
    const T* data() const { return storage_.data(); }
    iterator begin() { return storage_.begin(); }
    iterator end() { return storage_.end(); }
    const_iterator begin() const { return storage_.cbegin(); }
    const_iterator end() const { return storage_.cend(); }
    index_type iter_indices(const_iterator it) const {
        return dimension_.dimwise_indices(static_cast<size_type>(it - storage_.cbegin()));
    }

   private:
    MeshDimension<D> dimension_;
    std::vector<T> storage_;
};

template <typename T, std::size_t D, std::size_t O, typename U = double>
class InterpolationFunctionTemplate;

// ---- InterpolationFunction -----------------------------------------------

template <typename T, std::size_t D, std::size_t O, typename U = double>
class InterpolationFunction {
    static_assert(std::is_same<T, double>::value && std::is_same<U, double>::value,
                  "libintp_amd evaluates FP64 splines: T and U must be double");
    static_assert(D >= 1 && D <= INTP_MAX_DIM, "dimension without compiled kernels");
    static_assert(O <= INTP_MAX_ORDER, "order without compiled kernels");

   public:
    using val_type = T;
    using size_type = std::size_t;
    using coord_type = U;
    using diff_type = std::ptrdiff_t;
    constexpr static size_type dim = D;
    constexpr static size_type order = O;
    using function_type = InterpolationFunction<T, D, O, U>;
    template <typename V>
    using DimArray = std::array<V, D>;

    friend class InterpolationFunctionTemplate<T, D, O, U>;

    // 1D: (periodic, values range, (xmin, xmax)), Interpolation.hpp:49-79
    template <typename InputIter, typename C1, typename C2,
              typename std::enable_if<D == 1 && detail::is_iterator<InputIter>::value>::type* = nullptr>
    InterpolationFunction(bool periodic, std::pair<InputIter, InputIter> f_range, std::pair<C1, C2> x_range)
        : InterpolationFunction(DimArray<bool>{periodic}, Mesh<T, 1>(f_range),
                                std::make_pair(static_cast<typename std::common_type<C1, C2>::type>(x_range.first),
                                               static_cast<typename std::common_type<C1, C2>::type>(x_range.second))) {}

    template <typename InputIter, typename C1, typename C2,
              typename std::enable_if<D == 1 && detail::is_iterator<InputIter>::value>::type* = nullptr>
    InterpolationFunction(std::pair<InputIter, InputIter> f_range, std::pair<C1, C2> x_range)
        : InterpolationFunction(false, f_range, x_range) {}

    // (periodicity, mesh, one coordinate pair per dimension), :92-101
    template <typename... Ts>
    InterpolationFunction(DimArray<bool> periodicity, const Mesh<T, D>& f_mesh, std::pair<Ts, Ts>... x_ranges) {
        static_assert(sizeof...(Ts) == D, "one coordinate range per dimension");
        intp_grid_desc desc = describe(periodicity, f_mesh.dimension());
        std::vector<std::vector<double>> keep(D);
        detail::describe_all<D>(desc, keep, std::index_sequence_for<Ts...>{}, x_ranges...);
        intp_spline s = nullptr;
        detail::check(intp_build(detail::context(), &desc, f_mesh.data(), INTP_MEM_HOST, &s, nullptr));
        adopt(detail::own(s));
    }

    // all dimensions clamped, :104-107
    template <typename... Ts>
    InterpolationFunction(const Mesh<T, D>& f_mesh, std::pair<Ts, Ts>... x_ranges)
        : InterpolationFunction(DimArray<bool>{}, f_mesh, x_ranges...) {}

    // empty; populated by assignment from a built function, :125
    InterpolationFunction() = default;

    // ---- evaluation (Interpolation.hpp:132-244) ------------------------
    template <typename... Coords, typename = typename std::enable_if<
                                      std::is_arithmetic<typename std::common_type<Coords...>::type>::value>::type>
    val_type operator()(Coords... x) const {
        static_assert(sizeof...(Coords) == D, "one coordinate per dimension");
        return eval_one(DimArray<coord_type>{static_cast<coord_type>(x)...}, DimArray<size_type>{}, false);
    }
    val_type operator()(DimArray<coord_type> coord) const { return eval_one(coord, DimArray<size_type>{}, false); }

    val_type at(DimArray<coord_type> coord) const { return eval_one(coord, DimArray<size_type>{}, true); }
    template <typename... Coords, typename = typename std::enable_if<
                                      std::is_arithmetic<typename std::common_type<Coords...>::type>::value>::type>
    val_type at(Coords... x) const {
        return at(DimArray<coord_type>{static_cast<coord_type>(x)...});
    }

    val_type derivative(DimArray<coord_type> coord, DimArray<size_type> derivatives) const {
        return eval_one(coord, derivatives, false);
    }
    template <typename... Args>
    val_type derivative(DimArray<coord_type> coord, Args... deri_order) const {
        return derivative(coord, DimArray<size_type>{static_cast<size_type>(deri_order)...});
    }
    template <typename... CoordDeriOrderPair>
    val_type derivative(CoordDeriOrderPair... coord_deri_order) const {
        return derivative(DimArray<coord_type>{static_cast<coord_type>(coord_deri_order.first)...},
                          DimArray<size_type>{static_cast<size_type>(coord_deri_order.second)...});
    }

    val_type derivative_at(DimArray<coord_type> coord, DimArray<size_type> derivatives) const {
        return eval_one(coord, derivatives, true);
    }
    template <typename... Args>
    val_type derivative_at(DimArray<coord_type> coord, Args... deri_order) const {
        return derivative_at(coord, DimArray<size_type>{static_cast<size_type>(deri_order)...});
    }
    template <typename... CoordDeriOrderPair>
    val_type derivative_at(CoordDeriOrderPair... coord_deri_order) const {
        return derivative_at(DimArray<coord_type>{static_cast<coord_type>(coord_deri_order.first)...},
                             DimArray<size_type>{static_cast<size_type>(coord_deri_order.second)...});
    }

    // ---- batched evaluation (extension) --------------------------------
    // npts points, coordinate d of point i at coords[i * D + d] (AoS);
    // output o of point i at out[o * npts + i]; derivative orders
    // orders[o * D + d] (nullptr: the value only).  bounds_check: the at()
    // semantics.  Host buffers; one synchronous call.
    void eval(size_type npts, const coord_type* coords, val_type* out, size_type nout = 1,
              const size_type* orders = nullptr, bool bounds_check = false) const {
        eval_raw(npts, coords, 1, static_cast<std::int64_t>(D), out, static_cast<std::int64_t>(npts), 1, nout,
                 orders, INTP_EVAL_HOST_MEM | (bounds_check ? INTP_EVAL_BOUNDS_CHECK : 0u), nullptr);
    }
    // Device buffers with explicit strides (see intp_eval), asynchronous
    // on `stream` (a hipStream_t) unless bounds_check is set.
    void eval_device(size_type npts, const coord_type* coords, std::int64_t cs_dim, std::int64_t cs_pt,
                     val_type* out, std::int64_t os_out, std::int64_t os_pt, size_type nout = 1,
                     const size_type* orders = nullptr, bool bounds_check = false, void* stream = nullptr) const {
        eval_raw(npts, coords, cs_dim, cs_pt, out, os_out, os_pt, nout, orders,
                 bounds_check ? INTP_EVAL_BOUNDS_CHECK : 0u, stream);
    }

    // ---- properties (Interpolation.hpp:248-267) ------------------------
    bool periodicity(size_type d) const { return info().periodic[d] != 0; }
    bool uniform(size_type d) const { return info().uniform[d] != 0; }
    const std::pair<coord_type, coord_type>& range(size_type d) const { return ranges_[d]; }
    static constexpr size_type get_order() { return order; }

    // Control points (C-order) and knot vectors, copied to the host.
    std::vector<val_type> control_points() const {
        const intp_spline_info& in = info();
        std::int64_t count = 1;
        for (size_type d = 0; d < D; ++d) count *= in.n[d];
        std::vector<val_type> out(static_cast<std::size_t>(count));
        detail::check(intp_spline_download_ctrl(handle(), out.data(), count));
        return out;
    }
    std::vector<coord_type> knots(size_type d) const {
        std::vector<coord_type> out(static_cast<std::size_t>(info().nknots[d]));
        detail::check(intp_spline_download_knots(handle(), static_cast<std::int32_t>(d), out.data(),
                                                 static_cast<std::int64_t>(out.size())));
        return out;
    }

    // Value at a fixed coordinate of any function on the same grid
    // (Interpolation.hpp:493-506): a callable taking a function_type.
    auto eval_proxy(DimArray<coord_type> coords) const {
        return [coords](const function_type& f) { return f(coords); };
    }

   private:
    detail::SplineHandle spline_;
    intp_spline_info info_{};
    DimArray<std::pair<coord_type, coord_type>> ranges_{};

    static intp_grid_desc describe(const DimArray<bool>& periodicity, const MeshDimension<D>& dims) {
        intp_grid_desc desc{};
        desc.dim = static_cast<std::int32_t>(D);
        desc.order = static_cast<std::int32_t>(O);
        for (size_type d = 0; d < D; ++d) {
            desc.periodic[d] = periodicity[d] ? 1 : 0;
            desc.n[d] = static_cast<std::int64_t>(dims.dim_size(d));
        }
        return desc;
    }

    void adopt(detail::SplineHandle s) {
        spline_ = std::move(s);
        detail::check(intp_spline_get_info(spline_.get(), &info_));
        for (size_type d = 0; d < D; ++d) ranges_[d] = std::make_pair(info_.range_lo[d], info_.range_hi[d]);
    }

    intp_spline handle() const {
        if (!spline_) throw std::runtime_error("empty InterpolationFunction");
        return spline_.get();
    }
    const intp_spline_info& info() const {
        (void)handle();
        return info_;
    }

    void eval_raw(size_type npts, const coord_type* coords, std::int64_t cs_dim, std::int64_t cs_pt,
                  val_type* out, std::int64_t os_out, std::int64_t os_pt, size_type nout, const size_type* orders,
                  std::uint32_t flags, void* stream) const {
        std::vector<std::int32_t> r(nout * D, 0);
        if (orders)
            for (size_type i = 0; i < nout * D; ++i) r[i] = static_cast<std::int32_t>(orders[i]);
        detail::check(intp_eval(handle(), static_cast<std::int64_t>(npts), coords, cs_dim, cs_pt,
                                static_cast<std::int32_t>(nout), r.data(), out, os_out, os_pt, flags, stream));
    }

    val_type eval_one(const DimArray<coord_type>& coord, const DimArray<size_type>& r, bool bounds) const {
        val_type v = 0;
        eval_raw(1, coord.data(), 1, static_cast<std::int64_t>(D), &v, 1, 1, 1, r.data(),
                 INTP_EVAL_HOST_MEM | (bounds ? INTP_EVAL_BOUNDS_CHECK : 0u), nullptr);
        return v;
    }
};

// ---- InterpolationFunctionTemplate -----------------------------------------

template <typename T, std::size_t D, std::size_t O, typename U>
class InterpolationFunctionTemplate {
   public:
    using function_type = InterpolationFunction<T, D, O, U>;
    using size_type = typename function_type::size_type;
    using coord_type = typename function_type::coord_type;
    using val_type = typename function_type::val_type;
    using diff_type = typename function_type::diff_type;
    constexpr static size_type dim = D;
    constexpr static size_type order = O;
    template <typename V>
    using DimArray = std::array<V, D>;
    using MeshDim = MeshDimension<D>;

    // knots and factored collocation matrices of a grid,
    // InterpolationTemplate.hpp:60-79
    template <typename... Ts>
    InterpolationFunctionTemplate(DimArray<bool> periodicity, MeshDim mesh_dimension, std::pair<Ts, Ts>... x_ranges)
        : dimension_(mesh_dimension) {
        static_assert(sizeof...(Ts) == D, "one coordinate range per dimension");
        intp_grid_desc desc = function_type::describe(periodicity, mesh_dimension);
        std::vector<std::vector<double>> keep(D);
        detail::describe_all<D>(desc, keep, std::index_sequence_for<Ts...>{}, x_ranges...);
        intp_template t = nullptr;
        detail::check(intp_template_create(detail::context(), &desc, &t));
        template_ = detail::TemplateHandle(t, [](intp_template p) { intp_template_destroy(p); });
    }

    // 1D forms, :88-108
    template <typename C1, typename C2>
    InterpolationFunctionTemplate(bool periodicity, size_type f_length, std::pair<C1, C2> x_range)
        : InterpolationFunctionTemplate(DimArray<bool>{periodicity}, MeshDim(f_length), x_range) {
        static_assert(D == 1, "You can only use this overload of constructor in 1D case.");
    }
    template <typename C1, typename C2>
    InterpolationFunctionTemplate(size_type f_length, std::pair<C1, C2> x_range)
        : InterpolationFunctionTemplate(false, f_length, x_range) {}

    // all dimensions clamped, :118-124
    template <typename... Ts>
    InterpolationFunctionTemplate(MeshDim mesh_dimension, std::pair<Ts, Ts>... x_ranges)
        : InterpolationFunctionTemplate(DimArray<bool>{}, mesh_dimension, x_ranges...) {}

    // control points of one value mesh (or 1D iterator range), :118-143
    function_type interpolate(const Mesh<T, D>& mesh) const {
        for (size_type d = 0; d < D; ++d)
            if (mesh.dim_size(d) != dimension_.dim_size(d))
                throw std::runtime_error("value mesh shape does not match the template");
        intp_spline s = nullptr;
        detail::check(intp_interpolate(template_.get(), mesh.data(), INTP_MEM_HOST, &s, nullptr));
        function_type f;
        f.adopt(detail::own(s));
        return f;
    }
    template <typename It, typename = typename std::enable_if<D == 1 && detail::is_iterator<It>::value>::type>
    function_type interpolate(std::pair<It, It> range) const {
        return interpolate(Mesh<T, 1>(range));
    }
    template <typename MeshOrIterPair>
    void interpolate(function_type& interp, MeshOrIterPair&& mesh_or_range) const {
        interp = interpolate(std::forward<MeshOrIterPair>(mesh_or_range));
    }
    // Device-resident values (C-order doubles), asynchronous on `stream`.
    function_type interpolate_device(const val_type* values, void* stream = nullptr) const {
        intp_spline s = nullptr;
        detail::check(intp_interpolate(template_.get(), values, INTP_MEM_DEVICE, &s, stream));
        function_type f;
        f.adopt(detail::own(s));
        return f;
    }

    // value at a fixed coordinate of any function interpolated on this
    // grid, :145-176
    template <typename... Coords, typename = typename std::enable_if<
                                      std::is_arithmetic<typename std::common_type<Coords...>::type>::value>::type>
    auto eval_proxy(Coords... x) const {
        return eval_proxy(DimArray<coord_type>{static_cast<coord_type>(x)...});
    }
    auto eval_proxy(DimArray<coord_type> coord) const {
        return [coord](const function_type& f) { return f(coord); };
    }
    using eval_proxy_t = decltype(std::declval<function_type>().eval_proxy(DimArray<coord_type>{}));

   private:
    MeshDim dimension_;
    detail::TemplateHandle template_;
};

// ---- 1D conveniences -------------------------------------------------------

// x range [0, len - 1], or [0, len] for a periodic spline (no dummy
// point), Interpolation.hpp:509-540
template <std::size_t O = std::size_t{3}, typename T = double, typename U = double>
class InterpolationFunction1D : public InterpolationFunction<T, std::size_t{1}, O, U> {
    using base = InterpolationFunction<T, std::size_t{1}, O, U>;

   public:
    template <typename InputIter>
    InterpolationFunction1D(std::pair<InputIter, InputIter> f_range, bool periodicity = false)
        : InterpolationFunction1D(
              std::make_pair(typename base::coord_type{},
                             static_cast<typename base::coord_type>(std::distance(f_range.first, f_range.second) -
                                                                    (periodicity ? 0 : 1))),
              f_range, periodicity) {}

    template <typename C1, typename C2, typename InputIter>
    InterpolationFunction1D(std::pair<C1, C2> x_range, std::pair<InputIter, InputIter> f_range,
                            bool periodicity = false)
        : base(periodicity, f_range, x_range) {}
};

// InterpolationTemplate.hpp:583-604
template <std::size_t O, typename T = double, typename U = double>
class InterpolationFunctionTemplate1D : public InterpolationFunctionTemplate<T, std::size_t{1}, O, U> {
    using base = InterpolationFunctionTemplate<T, std::size_t{1}, O, U>;

   public:
    InterpolationFunctionTemplate1D(typename base::size_type f_length, bool periodicity = false)
        : InterpolationFunctionTemplate1D(
              std::make_pair(typename base::coord_type{}, static_cast<typename base::coord_type>(f_length - 1)),
              f_length, periodicity) {}

    template <typename C1, typename C2>
    InterpolationFunctionTemplate1D(std::pair<C1, C2> x_range, typename base::size_type f_length,
                                    bool periodicity = false)
        : base(periodicity, f_length, x_range) {}
};

}  // namespace intp

#endif
