// The drop-in's Mesh / MeshDimension surface against the reference's
// semantics (src/include/Mesh.hpp): C-order indexing, dim_acc_size,
// iterators along one dimension (skip iterators) and their indices,
// resize, allocator-aware construction; and the util helpers
// (src/include/util.hpp).  Host-only: needs no GPU and no library.  Prints one line per check; exits non-zero on a failure.
#include <intp/Interpolation.hpp>

#include <cstdio>
#include <initializer_list>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace {

int failures = 0;

void expect(const char* name, bool ok) {
    if (!ok) ++failures;
    std::printf("%-44s %s\n", name, ok ? "ok" : "FAILED");
}

// an allocator distinct from std::allocator, to exercise the Alloc parameter
template <typename T>
struct CountingAlloc {
    using value_type = T;
    CountingAlloc() = default;
    template <typename U>
    CountingAlloc(const CountingAlloc<U>&) {}
    T* allocate(std::size_t n) { return std::allocator<T>().allocate(n); }
    void deallocate(T* p, std::size_t n) { std::allocator<T>().deallocate(p, n); }
    bool operator==(const CountingAlloc&) const { return true; }
    bool operator!=(const CountingAlloc&) const { return false; }
};

}  // namespace

// The reference's util-test.cpp assertions on the helpers user code can
// name (index sequences, pow, the iterable / indexed traits), plus the
// default-initializing allocator.
void util_checks() {
    namespace u = intp::util;
    expect("util::make_index_sequence", std::is_same<u::make_index_sequence<3>, u::index_sequence<0u, 1u, 2u>>::value);
    expect("util::make_index_sequence_for",
           std::is_same<u::make_index_sequence_for<int, char>, u::index_sequence<0u, 1u>>::value);
    static_assert(u::pow(2, 3u) == 8, "compile-time pow");
    expect("util::pow", u::pow(2, 3u) == 8 && u::pow(1.5, 0u) == 1.0 && u::pow(3.0, 2u) == 9.0);
    expect("util::is_iterable", u::is_iterable<std::vector<int>>::value && !u::is_iterable<double>::value);
    expect("util::is_indexed",
           u::is_indexed<std::vector<int>>::value && !u::is_indexed<std::initializer_list<int>>::value &&
               !u::is_indexed<double*>::value);
    expect("util::remove_cvref_t", std::is_same<u::remove_cvref_t<const int&>, int>::value);
    std::vector<double, u::default_init_allocator<double>> v(4, 2.5);
    v.resize(6);
    v[5] = 1.0;
    expect("util::default_init_allocator", v.size() == 6 && v[3] == 2.5 && v[5] == 1.0);
    bool thrown = false;
    try {
        INTP_ASSERT(v.size() == 5, "INTP_ASSERT fired");
    } catch (const std::runtime_error&) {
        thrown = true;
    }
#ifdef INTP_ENABLE_ASSERTION
    expect("INTP_ASSERT throws with assertions on", thrown);
#else
    expect("INTP_ASSERT is empty with assertions off", !thrown);
#endif
}

int main() {
    util_checks();
    intp::MeshDimension<3> md(4, 5, 6);
    expect("size", md.size() == 120);
    expect("dim_acc_size", md.dim_acc_size(0) == 1 && md.dim_acc_size(1) == 6 && md.dim_acc_size(2) == 30 &&
                               md.dim_acc_size(3) == 120);
    expect("indexing (C order)", md.indexing({1, 2, 3}) == (1 * 5 + 2) * 6 + 3);
    expect("indexing_safe", md.indexing_safe(1, 2, 3) == md.indexing({1, 2, 3}));
    const auto ind = md.dimwise_indices(97);
    expect("dimwise_indices", md.indexing(ind) == 97);

    intp::Mesh<double, 3> m(4, 5, 6);
    std::iota(m.begin(), m.end(), 0.0);
    expect("operator()(i, j, k)", m(1, 2, 3) == static_cast<double>((1 * 5 + 2) * 6 + 3));
    // along dimension 1 through (2, *, 4): flat indices (2*5 + j)*6 + 4
    std::vector<double> line;
    for (auto it = m.begin(1, {2, 0, 4}); it != m.end(1, {2, 0, 4}); ++it) line.push_back(*it);
    bool ok = line.size() == 5;
    for (std::size_t j = 0; ok && j < 5; ++j) ok = line[j] == static_cast<double>((2 * 5 + j) * 6 + 4);
    expect("skip iterator along dim 1", ok);
    auto it0 = m.begin(0, {0, 3, 1});
    expect("skip iterator arithmetic", it0[2] == m(2, 3, 1) && (m.end(0, {0, 3, 1}) - it0) == 4 &&
                                           *(it0 + 3) == m(3, 3, 1));
    const auto& cm = m;
    auto cit = cm.begin(2, {1, 1, 0});
    ++cit;
    const auto at = cm.iter_indices(cit);
    expect("iter_indices of a skip iterator", at[0] == 1 && at[1] == 1 && at[2] == 1);
    *m.begin(2, {3, 4, 0}) = -1.0;
    expect("write through a skip iterator", m(3, 4, 0) == -1.0);

    intp::Mesh<double, 3, CountingAlloc<double>> a(intp::MeshDimension<3>(2, 3, 4));
    std::iota(a.begin(), a.end(), 1.0);
    intp::Mesh<double, 3> b(a);
    expect("mesh from another allocator", b.size() == 24 && b(1, 2, 3) == 24.0 && b(0, 0, 0) == 1.0);
    b.resize({3, 3, 3});
    expect("resize", b.size() == 27 && b.dim_size(0) == 3 && b.dimension().size() == 27);
    std::vector<double> v{1, 2, 3, 4};
    intp::Mesh<double, 1> r(intp::util::get_range(v));
    expect("1D mesh from an iterator range", r.size() == 4 && r(2) == 3.0);
#ifdef INTP_DEBUG
    // the reference's INTP_ASSERT range check (util.hpp:247-258)
    bool threw = false;
    try {
        (void)md.indexing_safe(4, 0, 0);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect("indexing_safe range check (INTP_DEBUG)", threw);
#endif
    std::printf("%d mesh check(s) failed\n", failures);
    return failures == 0 ? 0 : 1;
}
