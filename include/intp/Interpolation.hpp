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
//    spline / get_order, :248-267; eval_proxy, :493-506;
//  * intp::BSpline<T, D, O, U> from explicit knots and control points
//    (BSpline.hpp:188-216) and as a read-only view (BSpline.hpp:32-760);
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
// Conventions: the reference's periodic value layouts, chosen by the same
// macro -- with INTP_PERIODIC_NO_DUMMY_POINT a periodic dimension of a
// value mesh holds n values (spacing (max - min) / n), without it n + 1
// whose last, the dummy point, is ignored -- and INTP_CELL_LAYOUT
// semantics.
// Values: double (the FP64 path), float (the refined-FP32 form), or a
// fixed-size vector of either (component splines on one grid);
// coordinates double or float.
//
// Scalar members (operator(), at, derivative, derivative_at, eval_proxy)
// evaluate on the calling thread from a host copy of the GPU-built
// control points, made on first use, like the reference's per-point
// calls.  Added: batched evaluation of many points in one call (eval(),
// host or device buffers) on the GPU, which is how the GPU is meant to be
// used.
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
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../intp_amd.h"

// The reference's assertion switch and macro (util.hpp:247-258): with
// INTP_ENABLE_ASSERTION (implied by INTP_DEBUG), INTP_ASSERT(cond, msg)
// throws std::runtime_error(msg) when cond is false; otherwise it is empty.
#if defined(INTP_DEBUG) && !defined(INTP_ENABLE_ASSERTION)
#define INTP_ENABLE_ASSERTION
#endif
#ifndef INTP_ASSERT
#ifdef INTP_ENABLE_ASSERTION
#define INTP_ASSERT(assertion, msg)                                 \
    do {                                                            \
        if (!(assertion)) throw std::runtime_error(msg);            \
    } while (0)
#else
#define INTP_ASSERT(assertion, msg)
#endif
#endif

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

// The helpers of the reference's util.hpp that user code can name
// (util.hpp:18-48, 150-198, 278-303), on the standard library.
template <std::size_t... I>
using index_sequence = std::index_sequence<I...>;
template <std::size_t N>
using make_index_sequence = std::make_index_sequence<N>;
template <typename... T>
using make_index_sequence_for = std::index_sequence_for<T...>;

// base^exp for an unsigned exponent, multiplied as base * (base * (... * 1)).
template <typename T1, typename T2>
constexpr typename std::enable_if<std::is_unsigned<T2>::value, T1>::type pow(T1 base, T2 exp) {
    T1 r{1};
    for (T2 i = 0; i < exp; ++i) r = base * r;
    return r;
}

template <typename T>
using remove_cvref_t = typename std::remove_cv<typename std::remove_reference<T>::type>::type;

namespace detail_traits {
template <typename It>
using input_iterator_like = std::is_convertible<typename std::iterator_traits<It>::iterator_category,
                                                std::input_iterator_tag>;
template <typename T, typename = void>
struct iterable : std::false_type {};
template <typename T>
struct iterable<T, typename std::enable_if<input_iterator_like<decltype(std::declval<T&>().begin())>::value &&
                                           input_iterator_like<decltype(std::declval<T&>().end())>::value>::type>
    : std::true_type {};
template <typename T, typename = void>
struct indexed : std::false_type {};
template <typename T>
struct indexed<T, decltype(void(std::declval<const T&>().operator[](0)))> : std::true_type {};
}  // namespace detail_traits

// T has begin() and end() returning input iterators.
template <typename T>
struct is_iterable {
    static constexpr bool value = detail_traits::iterable<T>::value;
};

// const T has a member operator[](integer).
template <typename T>
struct is_indexed {
    static constexpr bool value = detail_traits::indexed<T>::value;
};

// An allocator that default-initializes (leaves trivial elements
// uninitialized) where a container would value-initialize.
template <typename T, typename A = std::allocator<T>>
struct default_init_allocator : public A {
    template <typename U>
    struct rebind {
        using other = default_init_allocator<U, typename std::allocator_traits<A>::template rebind_alloc<U>>;
    };
    using A::A;
    template <typename U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <typename U, typename... Args>
    void construct(U* p, Args&&... args) {
        std::allocator_traits<A>::construct(static_cast<A&>(*this), p, std::forward<Args>(args)...);
    }
};

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
    // product of the sizes of the last dim_ind dimensions (Mesh.hpp:45-51)
    size_type dim_acc_size(size_type dim_ind) const {
        size_type s = 1;
        for (size_type d = 0; d < dim_ind; ++d) s *= dim_size_[D - d - 1];
        return s;
    }
    // flat C-order index (last index fastest), Mesh.hpp:87-95
    size_type indexing(const index_type& ind) const {
        size_type flat = 0;
        for (size_type d = 0; d < D; ++d) flat = flat * dim_size_[d] + ind[d];
        return flat;
    }
    // indexing with the reference's range check, active like the
    // reference's INTP_ASSERT (Mesh.hpp:64-85, util.hpp:247-258)
    template <typename... Indices>
    size_type indexing_safe(Indices... ind) const {
        return indexing_safe(index_type{static_cast<size_type>(ind)...});
    }
    size_type indexing_safe(const index_type& ind) const {
#ifdef INTP_ENABLE_ASSERTION
        for (size_type d = D; d-- > 0;)
            if (ind[d] >= dim_size_[d])
                throw std::runtime_error("Mesh access out of range at dim " + std::to_string(d) + ", index " +
                                         std::to_string(ind[d]) + " is out of [0, " +
                                         std::to_string(dim_size_[d] - 1) + "].");
#endif
        return indexing(ind);
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
template <typename T, std::size_t D, typename Alloc = std::allocator<T>>
class Mesh {
   public:
    using size_type = std::size_t;
    using val_type = T;
    using allocator_type = Alloc;
    using index_type = typename MeshDimension<D>::index_type;
    using iterator = typename std::vector<T, Alloc>::iterator;
    using const_iterator = typename std::vector<T, Alloc>::const_iterator;
    constexpr static size_type dim = D;

    // Iterator along one dimension of the mesh (a fixed stride through the
    // storage), Mesh.hpp:138-236.
    template <typename V>
    class skip_iterator {
       public:
        using value_type = V;
        using difference_type = std::ptrdiff_t;
        using pointer = V*;
        using reference = V&;
        using iterator_category = std::random_access_iterator_tag;

        skip_iterator(pointer ptr, difference_type step) : ptr_(ptr), step_(step) {}
        operator skip_iterator<const V>() const { return {ptr_, step_}; }
        explicit operator const V*() const { return ptr_; }

        reference operator*() const { return *ptr_; }
        pointer operator->() const { return ptr_; }
        reference operator[](difference_type n) const { return *(ptr_ + n * step_); }
        bool operator==(const skip_iterator& o) const { return ptr_ == o.ptr_ && step_ == o.step_; }
        bool operator!=(const skip_iterator& o) const { return !(*this == o); }
        skip_iterator& operator++() {
            ptr_ += step_;
            return *this;
        }
        skip_iterator operator++(int) {
            skip_iterator t(*this);
            ++*this;
            return t;
        }
        skip_iterator& operator--() {
            ptr_ -= step_;
            return *this;
        }
        skip_iterator operator--(int) {
            skip_iterator t(*this);
            --*this;
            return t;
        }
        skip_iterator& operator+=(difference_type n) {
            ptr_ += n * step_;
            return *this;
        }
        skip_iterator& operator-=(difference_type n) {
            ptr_ -= n * step_;
            return *this;
        }
        skip_iterator operator+(difference_type n) const { return skip_iterator(*this) += n; }
        friend skip_iterator operator+(difference_type n, skip_iterator it) { return it += n; }
        skip_iterator operator-(difference_type n) const { return skip_iterator(*this) -= n; }
        difference_type operator-(const skip_iterator& o) const { return (ptr_ - o.ptr_) / step_; }
        bool operator<(const skip_iterator& o) const { return o - *this > 0; }
        bool operator>(const skip_iterator& o) const { return o < *this; }
        bool operator<=(const skip_iterator& o) const { return !(*this > o); }
        bool operator>=(const skip_iterator& o) const { return !(*this < o); }

       private:
        pointer ptr_;
        difference_type step_;
    };

    explicit Mesh(MeshDimension<D> dimension, const Alloc& alloc = Alloc())
        : dimension_(dimension), storage_(dimension.size(), T{}, alloc) {}
    explicit Mesh(index_type sizes) : Mesh(MeshDimension<D>(sizes)) {}
    template <typename... Args, typename = typename std::enable_if<(sizeof...(Args) == D && D > 1)>::type>
    explicit Mesh(Args... sizes) : Mesh(MeshDimension<D>(index_type{static_cast<size_type>(sizes)...})) {}
    explicit Mesh(size_type n, const Alloc& alloc = Alloc()) : Mesh(MeshDimension<D>(n), alloc) {}
    // 1D mesh from an iterator range
    template <typename It, typename = typename std::enable_if<D == 1 && detail::is_iterator<It>::value>::type>
    explicit Mesh(std::pair<It, It> range, const Alloc& alloc = Alloc())
        : dimension_(index_type{0}), storage_(alloc) {
        for (It it = range.first; it != range.second; ++it) storage_.push_back(static_cast<T>(*it));
        dimension_.resize(index_type{storage_.size()});
    }
    // a mesh with another allocator (Mesh.hpp:276-283)
    template <typename A2>
    Mesh(const Mesh<T, D, A2>& other, const Alloc& alloc = Alloc()) : Mesh(other.dimension(), alloc) {
        std::copy(other.begin(), other.end(), storage_.begin());
    }

    size_type size() const { return storage_.size(); }
    size_type dim_size(size_type d) const { return dimension_.dim_size(d); }
    const MeshDimension<D>& dimension() const { return dimension_; }
    void resize(index_type sizes) {
        dimension_.resize(sizes);
        storage_.resize(dimension_.size());
    }

    template <typename... Indices>
    T& operator()(Indices... ind) {
        return storage_[dimension_.indexing_safe(ind...)];
    }
    template <typename... Indices>
    const T& operator()(Indices... ind) const {
        return storage_[dimension_.indexing_safe(ind...)];
    }
    T& operator()(index_type ind) { return storage_[dimension_.indexing(ind)]; }
    const T& operator()(index_type ind) const { return storage_[dimension_.indexing(ind)]; }

    T* data() { return storage_.data(); }
    const T* data() const { return storage_.data(); }
    iterator begin() { return storage_.begin(); }
    iterator end() { return storage_.end(); }
    const_iterator begin() const { return storage_.cbegin(); }
    const_iterator end() const { return storage_.cend(); }
    // along dimension dim_ind through `indices` (its own entry ignored)
    skip_iterator<T> begin(size_type dim_ind, index_type indices) {
        indices[dim_ind] = 0;
        return {storage_.data() + dimension_.indexing_safe(indices), stride(dim_ind)};
    }
    skip_iterator<T> end(size_type dim_ind, index_type indices) {
        indices[dim_ind] = dimension_.dim_size(dim_ind);
        return {storage_.data() + dimension_.indexing(indices), stride(dim_ind)};
    }
    skip_iterator<const T> begin(size_type dim_ind, index_type indices) const {
        indices[dim_ind] = 0;
        return {storage_.data() + dimension_.indexing_safe(indices), stride(dim_ind)};
    }
    skip_iterator<const T> end(size_type dim_ind, index_type indices) const {
        indices[dim_ind] = dimension_.dim_size(dim_ind);
        return {storage_.data() + dimension_.indexing(indices), stride(dim_ind)};
    }
    index_type iter_indices(const_iterator it) const {
        return dimension_.dimwise_indices(static_cast<size_type>(it - storage_.cbegin()));
    }
    index_type iter_indices(skip_iterator<const T> it) const {
        return dimension_.dimwise_indices(static_cast<size_type>(static_cast<const T*>(it) - storage_.data()));
    }

   private:
    std::ptrdiff_t stride(size_type dim_ind) const {
        return static_cast<std::ptrdiff_t>(dimension_.dim_acc_size(D - dim_ind - 1));
    }

    MeshDimension<D> dimension_;
    std::vector<T, Alloc> storage_;
};

template <typename T, std::size_t D, std::size_t O, typename U = double>
class InterpolationFunctionTemplate;
template <typename T, std::size_t D, std::size_t O, typename U = double>
class InterpolationFunction;

namespace detail {

template <typename...>
using void_t = void;

// Value types: an arithmetic type (one component), or a fixed-size vector
// type with a static `dim`, a `value_type` and operator[] (the reference's
// own Vec<D, T>, test/src/include/Vec.hpp), interpolated component by
// component on one grid.  Components are float or double.
template <typename T, typename = void>
struct value_traits {
    static_assert(std::is_arithmetic<T>::value, "value type must be arithmetic or a fixed-size vector type");
    using scalar = T;
    static constexpr std::size_t components = 1;
    static scalar get(const T& v, std::size_t) { return v; }
    static void set(T& v, std::size_t, scalar x) { v = x; }
};

template <typename T>
struct value_traits<T, void_t<decltype(T::dim), typename T::value_type>> {
    using scalar = typename T::value_type;
    static constexpr std::size_t components = T::dim;
    static scalar get(const T& v, std::size_t c) { return v[c]; }
    static void set(T& v, std::size_t c, scalar x) { v[c] = x; }
};

// An evaluation proxy's device-library handle.
using ProxyHandle = std::shared_ptr<intp_proxy_s>;

inline ProxyHandle own(intp_proxy p) {
    return ProxyHandle(p, [](intp_proxy q) { intp_proxy_destroy(q); });
}

}  // namespace detail

// ---- BSpline -----------------------------------------------------------
//
// The reference's BSpline<T, D, O, U> (BSpline.hpp:32-760): built from
// explicit knot vectors and control points, or the read-only view of the
// spline inside an InterpolationFunction (InterpolationFunction::spline(),
// Interpolation.hpp:264): knot vectors, control points, range,
// periodicity and order, evaluation with or without knot-span hints, and
// pre_calc_coef.  Knots and control points are copied from the device on
// first access.  The loaders (load_knots, load_ctrlPts) build a new device
// table once everything is loaded; a table is never changed in place.
// Not provided: the basis helper base_spline_value and debug_output,
// which the reference's classes use internally.
// control_points() is the plain C-order table (the reference's
// INTP_CELL_LAYOUT build returns its redundant cell-layout copy instead).
template <typename T, std::size_t D, std::size_t O, typename U = double>
class BSpline {
    using traits = detail::value_traits<T>;
    static constexpr std::size_t C = traits::components;

   public:
    using val_type = T;
    using knot_type = U;
    using size_type = std::size_t;
    using diff_type = std::ptrdiff_t;
    constexpr static size_type dim = D;
    constexpr static size_type order = O;
    template <typename V>
    using DimArray = std::array<V, D>;
    using KnotContainer = std::vector<knot_type>;
    using knot_const_iterator = typename KnotContainer::const_iterator;
    using control_point_type = Mesh<val_type, D>;
    using ControlPointContainer = control_point_type;
    using spline_type = BSpline<T, D, O, U>;
    using BaseSpline = std::array<knot_type, O + 1>;

    BSpline() = default;

    // A spline from explicit knot vectors and control points
    // (BSpline.hpp:188-216): knot range d holds n[d] + order + 1 values,
    // plus order more when dimension d is periodic; an inconsistent count
    // throws std::runtime_error (the reference's INTP_ASSERT, :203-209).
    // The control points are copied to the device once, as FP64.
    template <typename... InputIters>
    BSpline(DimArray<bool> periodicity, const ControlPointContainer& ctrl_pts,
            std::pair<InputIters, InputIters>... knot_iter_pairs) {
        static_assert(sizeof...(InputIters) == D, "one knot range per dimension");
        build(periodicity, ctrl_pts,
              std::array<KnotContainer, D>{KnotContainer(knot_iter_pairs.first, knot_iter_pairs.second)...});
    }
    template <typename... InputIters>
    BSpline(const ControlPointContainer& ctrl_pts, std::pair<InputIters, InputIters>... knot_iter_pairs)
        : BSpline(DimArray<bool>{}, ctrl_pts, knot_iter_pairs...) {}

    // An empty spline of the given periodicity, filled by load_knots and
    // load_ctrlPts (BSpline.hpp:178-186, 217-242).  The device table is
    // built once every knot vector and the control points are loaded, and
    // again after each later load; range(d) follows load_knots at once.
    explicit BSpline(DimArray<bool> periodicity) : periodic_(periodicity) {}

    void load_knots(size_type d, KnotContainer knots) {
        if (d >= D) throw std::runtime_error("knot dimension out of range");
        if (knots.size() < 2 * O + 2) throw std::runtime_error("too few knots for the spline order");
        ranges_[d] = std::make_pair(knots[O], knots[knots.size() - O - (2 - O % 2)]);
        loaded_.knots[d] = std::move(knots);
        loaded_.have_knots[d] = true;
        build_loaded();
    }
    void load_ctrlPts(const ControlPointContainer& control_points) {
        loaded_.ctrl = control_points;
        loaded_.have_ctrl = true;
        build_loaded();
    }

    inline knot_const_iterator knots_begin(size_type d) const { return host().knots[d].cbegin(); }
    inline knot_const_iterator knots_end(size_type d) const { return host().knots[d].cend(); }
    inline size_type knots_num(size_type d) const { return host().knots[d].size(); }
    inline const control_point_type& control_points() const { return host().ctrl; }
    inline const std::pair<knot_type, knot_type>& range(size_type d) const { return ranges_[d]; }
    inline bool periodicity(size_type d) const { return periodic_[d]; }

   private:
    // The device table from knot vectors and control points
    // (intp_spline_from_knots; FP64 control points, one spline per value
    // component).
    void build(const DimArray<bool>& periodicity, const ControlPointContainer& ctrl_pts,
               const std::array<KnotContainer, D>& knots_in) {
        std::array<std::vector<double>, D> knots;
        for (size_type d = 0; d < D; ++d) knots[d].assign(knots_in[d].begin(), knots_in[d].end());
        std::array<std::int32_t, D> per{};
        std::array<std::int64_t, D> n{}, nk{};
        std::array<const double*, D> kp{};
        for (size_type d = 0; d < D; ++d) {
            per[d] = periodicity[d] ? 1 : 0;
            n[d] = static_cast<std::int64_t>(ctrl_pts.dim_size(d));
            nk[d] = static_cast<std::int64_t>(knots[d].size());
            kp[d] = knots[d].data();
        }
        std::array<detail::SplineHandle, C> splines;
        std::vector<double> comp;
        comp.reserve(ctrl_pts.size());
        for (std::size_t c = 0; c < C; ++c) {
            comp.clear();
            for (auto it = ctrl_pts.begin(); it != ctrl_pts.end(); ++it)
                comp.push_back(static_cast<double>(traits::get(*it, c)));
            intp_spline s = nullptr;
            detail::check(intp_spline_from_knots(detail::context(), static_cast<std::int32_t>(D),
                                                 static_cast<std::int32_t>(O), per.data(), n.data(), kp.data(),
                                                 nk.data(), comp.data(), &s));
            splines[c] = detail::own(s);
        }
        intp_spline_info info{};
        detail::check(intp_spline_get_info(splines[0].get(), &info));
        bind(splines, info);
    }
    void build_loaded() {
        if (!loaded_.have_ctrl) return;
        for (size_type d = 0; d < D; ++d)
            if (!loaded_.have_knots[d]) return;
        build(periodic_, loaded_.ctrl, loaded_.knots);
    }

   public:
    inline constexpr size_type get_order() const { return order; }

    // value at coordinates (BSpline.hpp:305-382)
    val_type operator()(DimArray<double> coords) const { return eval(coords, DimArray<size_type>{}, false); }
    template <typename... Coords, typename = typename std::enable_if<sizeof...(Coords) == D &&
                                                                     std::is_arithmetic<typename std::common_type<
                                                                         Coords...>::type>::value>::type>
    val_type operator()(Coords... x) const {
        return eval(DimArray<double>{static_cast<double>(x)...}, DimArray<size_type>{}, false);
    }
    // (coordinate, knot-span hint) per dimension (BSpline.hpp:305-368).  The
    // reference tries the hint before a binary search and finds the same
    // span either way; the span search here takes no hint.
    val_type operator()(DimArray<std::pair<knot_type, size_type>> coord_with_hints) const {
        DimArray<double> x;
        for (size_type d = 0; d < D; ++d) x[d] = static_cast<double>(coord_with_hints[d].first);
        return eval(x, DimArray<size_type>{}, false);
    }
    // mixed partial derivative, (coordinate, order) per dimension
    // (BSpline.hpp:540-550)
    val_type derivative_at(DimArray<std::pair<knot_type, size_type>> coord_deriv) const {
        DimArray<double> x;
        DimArray<size_type> r;
        for (size_type d = 0; d < D; ++d) {
            x[d] = static_cast<double>(coord_deriv[d].first);
            r[d] = coord_deriv[d].second;
        }
        return eval(x, r, false);
    }
    // (coordinate, hint, order) per dimension (BSpline.hpp:393-531)
    val_type derivative_at(DimArray<std::tuple<knot_type, size_type, size_type>> coord_hint_deriv) const {
        DimArray<double> x;
        DimArray<size_type> r;
        for (size_type d = 0; d < D; ++d) {
            x[d] = static_cast<double>(std::get<0>(coord_hint_deriv[d]));
            r[d] = std::get<2>(coord_hint_deriv[d]);
        }
        return eval(x, r, false);
    }
    // Basis values and stencil offset at a fixed coordinate, computed once;
    // the callable applies them to the control points of any spline on
    // the same knots (BSpline.hpp:244-297): a host proxy (intp_proxy_create).
    auto pre_calc_coef(DimArray<std::pair<knot_type, size_type>> coord_with_hints) const {
        DimArray<double> x;
        for (size_type d = 0; d < D; ++d) x[d] = static_cast<double>(coord_with_hints[d].first);
        intp_proxy p = nullptr;
        detail::check(intp_proxy_create(nullptr, handle(), 1, x.data(), 1, static_cast<std::int64_t>(D),
                                        INTP_EVAL_HOST_MEM, &p, nullptr));
        const detail::ProxyHandle ph = detail::own(p);
        return [ph](const spline_type& spline) {
            val_type v{};
            for (std::size_t c = 0; c < C; ++c) {
                double out = 0;
                detail::check(intp_proxy_eval(ph.get(), spline.handle(c), &out, 1, 0u, nullptr));
                traits::set(v, c, static_cast<typename traits::scalar>(out));
            }
            return v;
        };
    }

    // ---- used by InterpolationFunction ---------------------------------
    void bind(const std::array<detail::SplineHandle, C>& splines, const intp_spline_info& info) {
        splines_ = splines;
        info_ = info;
        for (size_type d = 0; d < D; ++d) periodic_[d] = info.periodic[d] != 0;
        for (size_type d = 0; d < D; ++d)
            ranges_[d] = std::make_pair(static_cast<knot_type>(info.range_lo[d]), static_cast<knot_type>(info.range_hi[d]));
        host_ = std::make_shared<HostCopy>();
    }
    // One point, derivative orders r, on the host copy of the table
    // (INTP_EVAL_HOST_MIRROR); `bounds`: the at() check.
    val_type eval(const DimArray<double>& x, const DimArray<size_type>& r, bool bounds) const {
        std::array<std::int32_t, D> ri{};
        for (size_type d = 0; d < D; ++d) ri[d] = static_cast<std::int32_t>(r[d]);
        val_type v{};
        for (std::size_t c = 0; c < C; ++c) {
            double out = 0;
            detail::check(intp_eval(handle(c), 1, x.data(), 1, static_cast<std::int64_t>(D), 1, ri.data(), &out, 1, 1,
                                    INTP_EVAL_HOST_MIRROR | (bounds ? INTP_EVAL_BOUNDS_CHECK : 0u), nullptr));
            traits::set(v, c, static_cast<typename traits::scalar>(out));
        }
        return v;
    }
    intp_spline handle(std::size_t c = 0) const {
        if (!splines_[c]) throw std::runtime_error("empty InterpolationFunction");
        return splines_[c].get();
    }
    const intp_spline_info& info() const {
        (void)handle();
        return info_;
    }

   private:
    struct HostCopy {
        std::once_flag once;
        std::array<KnotContainer, D> knots;
        control_point_type ctrl{MeshDimension<D>()};
    };
    // what load_knots / load_ctrlPts have received
    struct Loaded {
        std::array<KnotContainer, D> knots{};
        DimArray<bool> have_knots{};
        control_point_type ctrl{MeshDimension<D>()};
        bool have_ctrl = false;
    };
    std::array<detail::SplineHandle, C> splines_{};
    intp_spline_info info_{};
    DimArray<bool> periodic_{};
    DimArray<std::pair<knot_type, knot_type>> ranges_{};
    std::shared_ptr<HostCopy> host_;
    Loaded loaded_;

    const HostCopy& host() const {
        (void)handle();
        std::call_once(host_->once, [this]() {
            for (size_type d = 0; d < D; ++d) {
                std::vector<double> k(static_cast<std::size_t>(info_.nknots[d]));
                detail::check(intp_spline_download_knots(handle(), static_cast<std::int32_t>(d), k.data(),
                                                         static_cast<std::int64_t>(k.size())));
                host_->knots[d].assign(k.begin(), k.end());
            }
            typename MeshDimension<D>::index_type sizes{};
            std::int64_t count = 1;
            for (size_type d = 0; d < D; ++d) {
                sizes[d] = static_cast<std::size_t>(info_.n[d]);
                count *= info_.n[d];
            }
            control_point_type mesh{MeshDimension<D>(sizes)};
            std::vector<double> comp(static_cast<std::size_t>(count));
            for (std::size_t c = 0; c < C; ++c) {
                detail::check(intp_spline_download_ctrl(handle(c), comp.data(), count));
                auto it = mesh.begin();
                for (std::int64_t i = 0; i < count; ++i, ++it)
                    traits::set(*it, c, static_cast<typename traits::scalar>(comp[static_cast<std::size_t>(i)]));
            }
            host_->ctrl = std::move(mesh);
        });
        return *host_;
    }
};

// ---- InterpolationFunction -----------------------------------------------

template <typename T, std::size_t D, std::size_t O, typename U>
class InterpolationFunction {
    using traits = detail::value_traits<T>;
    using scalar_type = typename traits::scalar;
    static constexpr std::size_t C = traits::components;
    static_assert(std::is_same<scalar_type, double>::value || std::is_same<scalar_type, float>::value,
                  "values (or their components) must be double or float");
    static_assert(std::is_same<U, double>::value || std::is_same<U, float>::value,
                  "coordinates must be double or float");
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
    using spline_type = BSpline<T, D, O, U>;
    template <typename V>
    using DimArray = std::array<V, D>;

    friend class InterpolationFunctionTemplate<T, D, O, U>;

    // FP32 values (or components) are built and evaluated in the
    // refined-FP32 form: FP32 control points solved with FP64-residual
    // refinement (intp_build_refined, this many sweeps).
    static constexpr std::int32_t refine_sweeps = 1;

    // The periodic value layout (Interpolation.hpp:329-334, 379-388;
    // InterpolationTemplate.hpp:255-265, 463-486): without
    // INTP_PERIODIC_NO_DUMMY_POINT a periodic dimension of a value mesh
    // holds n + 1 values, the last (the dummy point at the period's end)
    // ignored; with it, n values.  The grid, and the library below, always
    // take the n values.
#ifdef INTP_PERIODIC_NO_DUMMY_POINT
    static constexpr bool dummy_point = false;
#else
    static constexpr bool dummy_point = true;
#endif

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
        std::array<detail::SplineHandle, C> splines;
        if (C == 1) {
            const std::vector<double> values = component(f_mesh, 0, periodicity);
            intp_spline s = nullptr;
            detail::check(refined() ? intp_build_refined(detail::context(), &desc, values.data(), INTP_MEM_HOST,
                                                         refine_sweeps, &s, nullptr)
                                    : intp_build(detail::context(), &desc, values.data(), INTP_MEM_HOST, &s, nullptr));
            splines[0] = detail::own(s);
        } else {
            // components share one grid: one template, one solve each
            intp_template t = nullptr;
            detail::check(intp_template_create(detail::context(), &desc, &t));
            const detail::TemplateHandle th(t, [](intp_template p) { intp_template_destroy(p); });
            for (std::size_t c = 0; c < C; ++c) splines[c] = interpolate_component(th.get(), f_mesh, c, periodicity);
        }
        adopt(splines);
    }

    // all dimensions clamped, :104-107
    template <typename... Ts>
    InterpolationFunction(const Mesh<T, D>& f_mesh, std::pair<Ts, Ts>... x_ranges)
        : InterpolationFunction(DimArray<bool>{}, f_mesh, x_ranges...) {}

    // empty; populated by assignment from a built function, :125
    InterpolationFunction() = default;

    // ---- evaluation (Interpolation.hpp:132-244) ------------------------
    // The scalar members evaluate on the calling thread from a host copy
    // of the control points (made on the first call), like the reference.
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

    // ---- batched evaluation (extension; scalar double values) ----------
    // npts points, coordinate d of point i at coords[i * D + d] (AoS);
    // output o of point i at out[o * npts + i]; derivative orders
    // orders[o * D + d] (nullptr: the value only).  bounds_check: the at()
    // semantics.  Host buffers; one synchronous call on the GPU.
    void eval(size_type npts, const double* coords, double* out, size_type nout = 1,
              const size_type* orders = nullptr, bool bounds_check = false) const {
        eval_raw(npts, coords, 1, static_cast<std::int64_t>(D), out, static_cast<std::int64_t>(npts), 1, nout,
                 orders, INTP_EVAL_HOST_MEM | (bounds_check ? INTP_EVAL_BOUNDS_CHECK : 0u), nullptr);
    }
    // Device buffers with explicit strides (see intp_eval), asynchronous
    // on `stream` (a hipStream_t) unless bounds_check is set.
    void eval_device(size_type npts, const double* coords, std::int64_t cs_dim, std::int64_t cs_pt, double* out,
                     std::int64_t os_out, std::int64_t os_pt, size_type nout = 1, const size_type* orders = nullptr,
                     bool bounds_check = false, void* stream = nullptr) const {
        eval_raw(npts, coords, cs_dim, cs_pt, out, os_out, os_pt, nout, orders,
                 bounds_check ? INTP_EVAL_BOUNDS_CHECK : 0u, stream);
    }

    // ---- properties (Interpolation.hpp:248-267) ------------------------
    bool periodicity(size_type d) const { return spline_.info().periodic[d] != 0; }
    bool uniform(size_type d) const { return spline_.info().uniform[d] != 0; }
    const std::pair<coord_type, coord_type>& range(size_type d) const {
        (void)spline_.handle();
        return spline_.range(d);
    }
    const spline_type& spline() const {
        (void)spline_.handle();
        return spline_;
    }
    static constexpr size_type get_order() { return order; }

    // Control points (C-order, scalar values: component 0) and knot
    // vectors, copied to the host.
    std::vector<double> control_points(std::size_t component = 0) const {
        const intp_spline_info& in = spline_.info();
        std::int64_t count = 1;
        for (size_type d = 0; d < D; ++d) count *= in.n[d];
        std::vector<double> out(static_cast<std::size_t>(count));
        detail::check(intp_spline_download_ctrl(spline_.handle(component), out.data(), count));
        return out;
    }
    std::vector<double> knots(size_type d) const {
        std::vector<double> out(static_cast<std::size_t>(spline_.info().nknots[d]));
        detail::check(intp_spline_download_knots(spline_.handle(), static_cast<std::int32_t>(d), out.data(),
                                                 static_cast<std::int64_t>(out.size())));
        return out;
    }

    // Value at a fixed coordinate of any function on the same grid
    // (Interpolation.hpp:493-506; BSpline.hpp:244-297): the knot spans and
    // basis values are computed here, once; the callable applies them to
    // a function's control points (a host proxy, intp_proxy_create).
    auto eval_proxy(DimArray<coord_type> coords) const {
        return make_proxy(nullptr, spline_.handle(), coords);
    }

    // the library handle of component c (extension)
    intp_spline handle(std::size_t c = 0) const { return spline_.handle(c); }

   private:
    spline_type spline_;

    static constexpr bool refined() { return std::is_same<scalar_type, float>::value; }

    static intp_grid_desc describe(const DimArray<bool>& periodicity, const MeshDimension<D>& dims) {
        intp_grid_desc desc{};
        desc.dim = static_cast<std::int32_t>(D);
        desc.order = static_cast<std::int32_t>(O);
        for (size_type d = 0; d < D; ++d) {
            desc.periodic[d] = periodicity[d] ? 1 : 0;
            const size_type n = dims.dim_size(d);
            desc.n[d] = static_cast<std::int64_t>(dummy_point && periodicity[d] && n > 0 ? n - 1 : n);
        }
        return desc;
    }

    // Component c of a value mesh, as doubles (C order), without the dummy
    // points of periodic dimensions.
    static std::vector<double> component(const Mesh<T, D>& mesh, std::size_t c, const DimArray<bool>& periodicity) {
        std::vector<double> v;
        v.reserve(mesh.size());
        DimArray<size_type> idx{};
        for (auto it = mesh.begin(); it != mesh.end(); ++it) {
            bool keep = true;
            if (dummy_point)
                for (size_type d = 0; d < D; ++d) keep = keep && !(periodicity[d] && idx[d] + 1 == mesh.dim_size(d));
            if (keep) v.push_back(static_cast<double>(traits::get(*it, c)));
            for (size_type d = D; d-- > 0;) {  // C order: the last index runs fastest
                if (++idx[d] < mesh.dim_size(d)) break;
                idx[d] = 0;
            }
        }
        return v;
    }

    static detail::SplineHandle interpolate_component(intp_template t, const Mesh<T, D>& mesh, std::size_t c,
                                                      const DimArray<bool>& periodicity) {
        const std::vector<double> values = component(mesh, c, periodicity);
        intp_spline s = nullptr;
        detail::check(refined() ? intp_interpolate_refined(t, values.data(), INTP_MEM_HOST, refine_sweeps, &s, nullptr)
                                : intp_interpolate(t, values.data(), INTP_MEM_HOST, &s, nullptr));
        return detail::own(s);
    }

    void adopt(const std::array<detail::SplineHandle, C>& splines) {
        intp_spline_info info{};
        detail::check(intp_spline_get_info(splines[0].get(), &info));
        spline_.bind(splines, info);
    }

    void eval_raw(size_type npts, const double* coords, std::int64_t cs_dim, std::int64_t cs_pt, double* out,
                  std::int64_t os_out, std::int64_t os_pt, size_type nout, const size_type* orders,
                  std::uint32_t flags, void* stream) const {
        static_assert(C == 1, "batched evaluation takes scalar values");
        std::vector<std::int32_t> r(nout * D, 0);
        if (orders)
            for (size_type i = 0; i < nout * D; ++i) r[i] = static_cast<std::int32_t>(orders[i]);
        detail::check(intp_eval(spline_.handle(), static_cast<std::int64_t>(npts), coords, cs_dim, cs_pt,
                                static_cast<std::int32_t>(nout), r.data(), out, os_out, os_pt, flags, stream));
    }

    val_type eval_one(const DimArray<coord_type>& coord, const DimArray<size_type>& r, bool bounds) const {
        DimArray<double> x;
        for (size_type d = 0; d < D; ++d) x[d] = static_cast<double>(coord[d]);
        return spline_.eval(x, r, bounds);
    }

    // A host proxy of one coordinate on the grid of template t or spline s.
    static auto make_proxy(intp_template t, intp_spline s, const DimArray<coord_type>& coord) {
        DimArray<double> x;
        for (size_type d = 0; d < D; ++d) x[d] = static_cast<double>(coord[d]);
        intp_proxy p = nullptr;
        detail::check(intp_proxy_create(t, s, 1, x.data(), 1, static_cast<std::int64_t>(D), INTP_EVAL_HOST_MEM, &p,
                                        nullptr));
        const detail::ProxyHandle ph = detail::own(p);
        return [ph](const function_type& f) {
            val_type v{};
            for (std::size_t c = 0; c < C; ++c) {
                double out = 0;
                detail::check(intp_proxy_eval(ph.get(), f.handle(c), &out, 1, 0u, nullptr));
                traits::set(v, c, static_cast<scalar_type>(out));
            }
            return v;
        };
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
    // Without INTP_PERIODIC_NO_DUMMY_POINT, mesh_dimension counts the dummy
    // point of each periodic dimension, as the meshes given to interpolate().
    InterpolationFunctionTemplate(DimArray<bool> periodicity, MeshDim mesh_dimension, std::pair<Ts, Ts>... x_ranges)
        : dimension_(mesh_dimension), periodicity_(periodicity) {
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
        std::array<detail::SplineHandle, detail::value_traits<T>::components> splines;
        for (std::size_t c = 0; c < splines.size(); ++c)
            splines[c] = function_type::interpolate_component(template_.get(), mesh, c, periodicity_);
        function_type f;
        f.adopt(splines);
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
    // Device-resident values (C-order doubles, scalar double splines,
    // n values per periodic dimension whatever the dummy-point
    // convention), asynchronous on `stream`.
    function_type interpolate_device(const double* values, void* stream = nullptr) const {
        static_assert(detail::value_traits<T>::components == 1, "device values are scalar");
        intp_spline s = nullptr;
        detail::check(intp_interpolate(template_.get(), values, INTP_MEM_DEVICE, &s, stream));
        function_type f;
        f.adopt({detail::own(s)});
        return f;
    }

    // value at a fixed coordinate of any function interpolated on this
    // grid, :145-176: spans and basis computed once (a host proxy)
    template <typename... Coords, typename = typename std::enable_if<
                                      std::is_arithmetic<typename std::common_type<Coords...>::type>::value>::type>
    auto eval_proxy(Coords... x) const {
        return eval_proxy(DimArray<coord_type>{static_cast<coord_type>(x)...});
    }
    auto eval_proxy(DimArray<coord_type> coord) const { return function_type::make_proxy(template_.get(), nullptr, coord); }
    using eval_proxy_t = decltype(std::declval<function_type>().eval_proxy(DimArray<coord_type>{}));

   private:
    MeshDim dimension_;
    DimArray<bool> periodicity_{};
    detail::TemplateHandle template_;
};

// ---- 1D conveniences -------------------------------------------------------

// x range [0, len - 1], or [0, len] for a periodic spline under
// INTP_PERIODIC_NO_DUMMY_POINT (Interpolation.hpp:509-540)
template <std::size_t O = std::size_t{3}, typename T = double, typename U = double>
class InterpolationFunction1D : public InterpolationFunction<T, std::size_t{1}, O, U> {
    using base = InterpolationFunction<T, std::size_t{1}, O, U>;

   public:
    template <typename InputIter>
    InterpolationFunction1D(std::pair<InputIter, InputIter> f_range, bool periodicity = false)
        : InterpolationFunction1D(
              std::make_pair(typename base::coord_type{},
                             static_cast<typename base::coord_type>(std::distance(f_range.first, f_range.second) -
                                                                    (periodicity && !base::dummy_point ? 0 : 1))),
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
