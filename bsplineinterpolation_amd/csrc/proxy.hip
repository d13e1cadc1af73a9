// proxy.hip -- evaluation proxies on the device: the knot spans and basis
// weights of a fixed set of points are computed once (the reference's
// pre_calc_coef, BSpline.hpp:244-297, for many points at a time), then
// applied to the control points of any spline on the same grid with one
// gather and contraction per point.  The weights are the evaluation
// kernels' own (per-class polynomials on uniform grids, the Cox-de Boor
// triangle otherwise) and the contraction is theirs, so a proxy's value
// equals intp_eval's value bit for bit.
#include "eval_impl.hpp"

namespace intp_amd {

namespace {

constexpr int kProxyBlock = 256;

// Span and value weights of one coordinate on a general grid (the first
// half of general_weights, keeping the span).
template <int K>
__device__ __forceinline__ int general_span_weights(const DimParams& dp, double x, int* bad, double (&w)[K + 1]) {
    wrap_or_check(dp, x, bad);
    const int last = dp.nt - K - 2;
    const int mu0 = dp.uniform ? hint<K>(dp, x, last) : K;
    const double* __restrict__ t = dp.knots;
    const int mu = span<K>(mu0, last, x, [t](int i) { return t[i]; });
    double R[(K * (K + 1)) / 2 + 1];
    double keep[K + 1];
    triangle<K>(dp, mu, x, w, keep, K, R);
    return mu;
}

template <int D, int K, bool UNIFORM>
__global__ __launch_bounds__(kProxyBlock) void proxy_create_kernel(ProxyParams p) {
    const int64_t pt = static_cast<int64_t>(blockIdx.x) * kProxyBlock + threadIdx.x;
    if (pt >= p.npts) return;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double x = p.coords[d * p.cs_dim + pt * p.cs_pt];
        double* w = p.weights + (pt * D + d) * (K + 1);
        int mu;
        if constexpr (UNIFORM) {
            DimW<K, uint32_t, double> dd;
            mu = uniform_weights<K, false, double>(p.dim[d], x, 0, p.bad, dd);
#pragma unroll
            for (int m = 0; m <= K; ++m) w[m] = dd.w[m];
        } else {
            double b[K + 1];
            mu = general_span_weights<K>(p.dim[d], x, p.bad, b);
#pragma unroll
            for (int m = 0; m <= K; ++m) w[m] = b[m];
        }
        p.first[pt * D + d] = mu - K;
    }
}

template <int D, int K, bool LASTC, typename T, typename Off>
__global__ __launch_bounds__(kProxyBlock) void proxy_apply_kernel(ProxyParams p) {
    const int64_t pt = static_cast<int64_t>(blockIdx.x) * kProxyBlock + threadIdx.x;
    if (pt >= p.npts) return;
    DimW<K, Off, T> dd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const DimParams& dp = p.dim[d];
        const double* w = p.weights + (pt * D + d) * (K + 1);
        const Off first = static_cast<Off>(p.first[pt * D + d]);
        const Off n = static_cast<Off>(dp.n), stride = static_cast<Off>(dp.stride);
#pragma unroll
        for (int m = 0; m <= K; ++m) {
            Off id = first + static_cast<Off>(m);
            if (dp.periodic && id >= n) id -= n;
            dd[d].off[m] = id * stride;
            dd[d].w[m] = static_cast<T>(w[m]);
            dd[d].dw[m] = T(0);
        }
    }
    const T v = contract_single<D, K, LASTC>(static_cast<const T*>(p.ctrl), dd);
    if (p.out32)
        p.out32[pt * p.os_pt] = static_cast<float>(v);
    else
        p.out[pt * p.os_pt] = static_cast<double>(v);
}

template <void (*KF)(ProxyParams)>
hipError_t launch_points(const ProxyParams& p, hipStream_t s) {
    const int64_t blocks = (p.npts + kProxyBlock - 1) / kProxyBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KF, dim3(static_cast<unsigned>(blocks)), dim3(kProxyBlock), 0, s, p);
    return hipGetLastError();
}

template <int D, int K>
hipError_t create_dk(bool uniform, const ProxyParams& p, hipStream_t s) {
    if (uniform) return launch_points<proxy_create_kernel<D, K, true>>(p, s);
    return launch_points<proxy_create_kernel<D, K, false>>(p, s);
}

template <int D, int K, typename T>
hipError_t apply_dkt(bool uniform, const ProxyParams& p, hipStream_t s) {
    // 32-bit offsets on the uniform path (tables below 2^32 elements), as
    // the evaluation kernels use
    const bool lastc = !p.dim[D - 1].periodic;
    if (uniform) {
        if (lastc) return launch_points<proxy_apply_kernel<D, K, true, T, uint32_t>>(p, s);
        return launch_points<proxy_apply_kernel<D, K, false, T, uint32_t>>(p, s);
    }
    if (lastc) return launch_points<proxy_apply_kernel<D, K, true, T, int64_t>>(p, s);
    return launch_points<proxy_apply_kernel<D, K, false, T, int64_t>>(p, s);
}

template <int D>
hipError_t create_d(int order, bool uniform, const ProxyParams& p, hipStream_t s) {
    switch (order) {
        case 0: return create_dk<D, 0>(uniform, p, s);
        case 1: return create_dk<D, 1>(uniform, p, s);
        case 2: return create_dk<D, 2>(uniform, p, s);
        case 3: return create_dk<D, 3>(uniform, p, s);
        case 4: return create_dk<D, 4>(uniform, p, s);
        case 5: return create_dk<D, 5>(uniform, p, s);
        case 6: return create_dk<D, 6>(uniform, p, s);
        case 7: return create_dk<D, 7>(uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

template <int D, typename T>
hipError_t apply_d(int order, bool uniform, const ProxyParams& p, hipStream_t s) {
    switch (order) {
        case 0: return apply_dkt<D, 0, T>(uniform, p, s);
        case 1: return apply_dkt<D, 1, T>(uniform, p, s);
        case 2: return apply_dkt<D, 2, T>(uniform, p, s);
        case 3: return apply_dkt<D, 3, T>(uniform, p, s);
        case 4: return apply_dkt<D, 4, T>(uniform, p, s);
        case 5: return apply_dkt<D, 5, T>(uniform, p, s);
        case 6: return apply_dkt<D, 6, T>(uniform, p, s);
        case 7: return apply_dkt<D, 7, T>(uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

__global__ __launch_bounds__(kProxyBlock) void first_outside_kernel(BoundsParams p) {
    const int64_t pt = static_cast<int64_t>(blockIdx.x) * kProxyBlock + threadIdx.x;
    if (pt >= p.npts) return;
    bool out = false;
    for (int d = 0; d < p.ndim; ++d) {
        const DimParams& dp = p.dim[d];
        const double x = p.coords[d * p.cs_dim + pt * p.cs_pt];
        out = out || (!dp.periodic && !(x >= dp.lo && x <= dp.hi));
    }
    if (out) atomicMin(reinterpret_cast<unsigned long long*>(p.first), static_cast<unsigned long long>(pt));
}

__global__ __launch_bounds__(kProxyBlock) void fill_u32_kernel(uint32_t* p, uint32_t value, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kProxyBlock + threadIdx.x;
    if (i < n) p[i] = value;
}

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, int64_t n, hipStream_t s) {
    const int64_t blocks = (n + kProxyBlock - 1) / kProxyBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kProxyBlock), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_first_outside(const BoundsParams& p, hipStream_t s) {
    const int64_t blocks = (p.npts + kProxyBlock - 1) / kProxyBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF || p.ndim < 1 || p.ndim > kMaxDim) return hipErrorInvalidValue;
    hipLaunchKernelGGL(first_outside_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kProxyBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_proxy_create(int dim, int order, bool uniform, const ProxyParams& p, hipStream_t s) {
    switch (dim) {
        case 1: return create_d<1>(order, uniform, p, s);
        case 2: return create_d<2>(order, uniform, p, s);
        case 3: return create_d<3>(order, uniform, p, s);
        case 4: return create_d<4>(order, uniform, p, s);
        case 5: return create_d<5>(order, uniform, p, s);
        case 6: return create_d<6>(order, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_proxy_apply(int dim, int order, bool uniform, bool f32, const ProxyParams& p, hipStream_t s) {
    switch (dim * 2 + (f32 ? 1 : 0)) {
        case 2: return apply_d<1, double>(order, uniform, p, s);
        case 3: return apply_d<1, float>(order, uniform, p, s);
        case 4: return apply_d<2, double>(order, uniform, p, s);
        case 5: return apply_d<2, float>(order, uniform, p, s);
        case 6: return apply_d<3, double>(order, uniform, p, s);
        case 7: return apply_d<3, float>(order, uniform, p, s);
        case 8: return apply_d<4, double>(order, uniform, p, s);
        case 9: return apply_d<4, float>(order, uniform, p, s);
        case 10: return apply_d<5, double>(order, uniform, p, s);
        case 11: return apply_d<5, float>(order, uniform, p, s);
        case 12: return apply_d<6, double>(order, uniform, p, s);
        case 13: return apply_d<6, float>(order, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace intp_amd
