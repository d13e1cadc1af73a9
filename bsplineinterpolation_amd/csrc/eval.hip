// eval.hip -- batched evaluation of tensor-product B-splines on gfx950.
//
// One lane per query point.  Per dimension, in registers: periodic wrap
// (BSpline.hpp:130-136) and the knot span from the uniform-grid hint with a
// binary-search fallback (Interpolation.hpp:277-295, BSpline.hpp:125-157),
// then the basis weights:
//  * uniform grids (fast path): knots are recomputed from their closed form
//    with the host's exact operation order (no FMA), so the span test is
//    the reference's, with no memory traffic; the basis comes from the
//    span's polynomial coefficients (one class for all interior spans,
//    one per clamped end span), staged in LDS;
//  * other grids: knots and reciprocal knot gaps are read from HBM and the
//    Cox-de Boor triangle (BSpline.hpp:83-111) runs in registers.
// Derivatives: the reference differentiates the gathered control points
// (BSpline.hpp:489-520); here the same difference operators act,
// transposed, on the basis weights (or the basis polynomials are
// differentiated), so all requested outputs share one gather of the
// (k+1)^D stencil.  The stencil is contracted one axis at a time, last
// (contiguous) axis first.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kBlock = 256;

#ifndef INTP_OUTER_UNROLL_MAX_K
#define INTP_OUTER_UNROLL_MAX_K 2  // fully unroll the outer stencil loop up to this order
#endif

template <int K, typename Off>
struct DimW {
    double w[K + 1];   // value weights (or the single output's weights)
    double dw[K + 1];  // first-derivative weights (gradient kernels)
    Off off[K + 1];    // element offsets of the K+1 stencil planes
};

// ---------------------------------------------------------------------------
// span search
// ---------------------------------------------------------------------------

__device__ __forceinline__ void wrap_or_check(const DimParams& dp, double& x, int* bad) {
    if (dp.periodic) {
        double r = x - dp.lo;
        if (!(r >= 0.0 && r < dp.period)) r = fmod(r, dp.period);
        x = dp.lo + r + (x < dp.lo ? dp.period : 0.0);
    } else if (bad != nullptr && (x < dp.lo || x > dp.hi)) {
        *bad = 1;
    }
}

template <int K>
__device__ __forceinline__ int hint(const DimParams& dp, double x, int last) {
    double v = (x - dp.lo) * dp.inv_dx - dp.shift;
    v = v > 0.0 ? v : 0.0;
    return v >= static_cast<double>(last - K) ? last : min(last, static_cast<int>(ceil(v)) + K);
}

// The unique mu in [K, last] with t[mu] <= x < t[mu+1], clamped (the
// result of BSpline.hpp:146-156); `T` yields knot i.
template <int K, typename KnotFn>
__device__ __forceinline__ int span(int mu, int last, double x, KnotFn T) {
    if (T(mu) <= x && x < T(mu + 1)) return mu;
    if (mu > K && T(mu - 1) <= x && x < T(mu)) return mu - 1;
    if (mu < last && T(mu + 1) <= x && x < T(mu + 2)) return mu + 1;
    int lo = K + 1, hi = last + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x < T(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo - 1;
}

// ---------------------------------------------------------------------------
// basis weights, general grids
// ---------------------------------------------------------------------------

// Cox-de Boor triangle on span mu.  `keep` receives level `keep_level`
// (the basis of order keep_level, right-aligned); `b` ends at level K.
template <int K>
__device__ __forceinline__ void triangle(const DimParams& dp, int mu, double x, double (&b)[K + 1],
                                         double (&keep)[K + 1], int keep_level,
                                         double (&R)[(K * (K + 1)) / 2 + 1]) {
    constexpr int T = (K * (K + 1)) / 2;
    const double* __restrict__ t = dp.knots;
    double tt[2 * K + 1];
#pragma unroll
    for (int j = 0; j < 2 * K; ++j) tt[j] = t[mu - K + 1 + j];  // t[mu-K+1 .. mu+K]
    const double* __restrict__ rr = dp.recip + static_cast<int64_t>(mu) * T;
#pragma unroll
    for (int j = 0; j < T; ++j) R[j] = rr[j];
#pragma unroll
    for (int m = 0; m <= K; ++m) b[m] = 0.0;
    b[K] = 1.0;
#pragma unroll
    for (int m = 0; m <= K; ++m) keep[m] = b[m];
#pragma unroll
    for (int i = 1; i <= K; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            const int idx = K - i + j;
            // left: b * (x - t[mu-i+j]) / (t[mu+j] - t[mu-i+j])
            const double left = j == 0 ? 0.0 : b[idx] * (x - tt[K - 1 - i + j]) * R[i * (i - 1) / 2 + j - 1];
            // right: b * (t[mu+j+1] - x) / (t[mu+j+1] - t[mu-i+j+1])
            const double right = idx == K ? 0.0 : b[idx + 1] * (tt[K + j] - x) * R[i * (i - 1) / 2 + j];
            b[idx] = left + right;
        }
        if (i == keep_level) {
#pragma unroll
            for (int m = 0; m <= K; ++m) keep[m] = b[m];
        }
    }
}

// Weights on the original control points of the r-th derivative, from the
// right-aligned order-(K-r) basis `w` (transpose of BSpline.hpp:507-518).
template <int K>
__device__ __forceinline__ void differentiate(double (&w)[K + 1], int r, const double (&R)[(K * (K + 1)) / 2 + 1]) {
#pragma unroll
    for (int kk = 1; kk <= K; ++kk) {
        if (kk > K - r) {
            double nw[K + 1];
            const int first = K - kk + 1;  // affected positions first..K
#pragma unroll
            for (int m = 0; m <= K; ++m) {
                double v;
                if (m >= first)
                    v = w[m] * (static_cast<double>(kk) * R[kk * (kk - 1) / 2 + (m - K + kk) - 1]);
                else
                    v = w[m];
                if (m + 1 >= first && m + 1 <= K)
                    v -= w[m + 1] * (static_cast<double>(kk) * R[kk * (kk - 1) / 2 + (m + 1 - K + kk) - 1]);
                nw[m] = v;
            }
#pragma unroll
            for (int m = 0; m <= K; ++m) w[m] = nw[m];
        }
    }
}

template <int K, typename Off>
__device__ __forceinline__ void stencil_offsets(const DimParams& dp, int mu, Off (&off)[K + 1]) {
    const Off base = static_cast<Off>(mu - K);
    const Off n = static_cast<Off>(dp.n), stride = static_cast<Off>(dp.stride);
#pragma unroll
    for (int m = 0; m <= K; ++m) {
        Off id = base + static_cast<Off>(m);
        if (dp.periodic && id >= n) id -= n;
        off[m] = id * stride;
    }
}

// General grids: weights of the value (w) and, if `grad`, of d/dx (dw);
// otherwise w holds the r-th derivative.
template <int K, bool GRAD>
__device__ __forceinline__ void general_weights(const DimParams& dp, double x, int r, int* bad,
                                                DimW<K, int64_t>& dd) {
    wrap_or_check(dp, x, bad);
    const int last = dp.nt - K - 2;
    const int mu0 = dp.uniform ? hint<K>(dp, x, last) : K;
    const double* __restrict__ t = dp.knots;
    const int mu = span<K>(mu0, last, x, [t](int i) { return t[i]; });
    double R[(K * (K + 1)) / 2 + 1];
    double keep[K + 1];
    double b[K + 1];
    triangle<K>(dp, mu, x, b, keep, GRAD ? K - 1 : K - r, R);
    if constexpr (GRAD) {
#pragma unroll
        for (int m = 0; m <= K; ++m) dd.w[m] = b[m];
        if (K >= 1) differentiate<K>(keep, 1, R);
#pragma unroll
        for (int m = 0; m <= K; ++m) dd.dw[m] = K >= 1 ? keep[m] : 0.0;
    } else {
        if (r == 0) {
#pragma unroll
            for (int m = 0; m <= K; ++m) keep[m] = b[m];
        } else if (r <= K) {
            differentiate<K>(keep, r, R);
        }
#pragma unroll
        for (int m = 0; m <= K; ++m) dd.w[m] = keep[m];
    }
    stencil_offsets<K>(dp, mu, dd.off);
}

// ---------------------------------------------------------------------------
// basis weights, uniform grids
// ---------------------------------------------------------------------------

// Knot i of a uniform dimension without branches (selects only).
template <int K>
__device__ __forceinline__ double uknot_sel(const DimParams& dp, int i) {
#pragma clang fp contract(off)
    const double f = dp.xmin + (static_cast<double>(i) - dp.half) * dp.dx;
    if (dp.periodic) return f;
    return i <= K ? dp.xmin : (i >= dp.nt - K - 1 ? dp.xmax : f);
}

// Span of x on a uniform dimension: hint, one select-based correction,
// then a check against the exact knots; a lane that still disagrees (NaN,
// coordinates beyond int range) takes the binary search.
template <int K>
__device__ __forceinline__ int uniform_span(const DimParams& dp, double x) {
    const int last = dp.nt - K - 2;
    double v = (x - dp.lo) * dp.inv_dx - dp.shift;
    v = v > 0.0 ? v : 0.0;
    v = v < static_cast<double>(last) ? v : static_cast<double>(last);
    int mu = min(last, static_cast<int>(ceil(v)) + K);
    const double tl = uknot_sel<K>(dp, mu), th = uknot_sel<K>(dp, mu + 1);
    mu += (x >= th && mu < last) ? 1 : ((x < tl && mu > K) ? -1 : 0);
    const bool ok = (mu == K || uknot_sel<K>(dp, mu) <= x) && (mu == last || x < uknot_sel<K>(dp, mu + 1));
    if (!ok) {
        int lo = K + 1, hi = last + 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (x < uknot_sel<K>(dp, mid))
                hi = mid;
            else
                lo = mid + 1;
        }
        mu = lo - 1;
    }
    return mu;
}

template <int K, bool GRAD>
__device__ __forceinline__ void uniform_weights(const DimParams& dp, const double* __restrict__ poly, double x,
                                                int r, int* bad, DimW<K, uint32_t>& dd) {
    wrap_or_check(dp, x, bad);
    const int last = dp.nt - K - 2;
    const int mu = uniform_span<K>(dp, x);
    const int cls = (mu - K < dp.nleft) ? 1 + (mu - K) : ((last - mu < dp.nright) ? 1 + dp.nleft + (last - mu) : 0);
    const double u = (x - uknot_sel<K>(dp, mu)) * dp.inv_dx;
    const double* __restrict__ P = poly + cls * (K + 1) * (K + 1);
    constexpr int W = K + 1;
    if constexpr (GRAD) {
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            double v = P[j * W + K];
            double g = static_cast<double>(K) * P[j * W + K];
#pragma unroll
            for (int p = K - 1; p >= 0; --p) {
                v = v * u + P[j * W + p];
                if (p >= 1) g = g * u + static_cast<double>(p) * P[j * W + p];
            }
            dd.w[j] = v;
            dd.dw[j] = K >= 1 ? g * dp.inv_dx : 0.0;
        }
    } else {
        // r-th derivative: sum_p p!/(p-r)! P_p u^(p-r), times dx^-r
        double scale = 1.0;
        for (int i = 0; i < r; ++i) scale *= dp.inv_dx;
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            double v = 0.0;
#pragma unroll
            for (int p = K; p >= 0; --p) {
                if (p >= r) {
                    double ff = 1.0;
#pragma unroll
                    for (int q = 0; q < K; ++q)
                        if (q < r) ff *= static_cast<double>(p - q);
                    v = v * u + ff * P[j * W + p];
                }
            }
            dd.w[j] = r > K ? 0.0 : v * scale;
        }
    }
    stencil_offsets<K>(dp, mu, dd.off);
}

// ---------------------------------------------------------------------------
// stencil contraction
// ---------------------------------------------------------------------------

// One stencil row along the last dimension.  LASTC: the last dimension is
// not periodic, so the row is K+1 consecutive elements from one address
// (one address computation, immediate offsets).
template <int K, bool LASTC, typename Off>
__device__ __forceinline__ void load_row(const double* __restrict__ c, Off rowoff, const DimW<K, Off>& dl,
                                         double (&cv)[K + 1]) {
    if constexpr (LASTC) {
        const double* __restrict__ r = c + (rowoff + dl.off[0]);
#pragma unroll
        for (int e = 0; e <= K; ++e) cv[e] = r[e];
    } else {
        const double* __restrict__ r = c + rowoff;
#pragma unroll
        for (int e = 0; e <= K; ++e) cv[e] = r[dl.off[e]];
    }
}

// Shifts the weights/offsets of the first dimension down by one, so a
// rolled loop over its K+1 planes only ever reads element 0 (registers,
// never scratch).
template <int K, typename Off>
__device__ __forceinline__ void shift_down(DimW<K, Off>& d) {
#pragma unroll
    for (int m = 0; m < K; ++m) {
        d.w[m] = d.w[m + 1];
        d.dw[m] = d.dw[m + 1];
        d.off[m] = d.off[m + 1];
    }
}

template <int D, int K, bool LASTC, typename Off>
__device__ __forceinline__ void contract_grad(const double* __restrict__ c, DimW<K, Off> (&dd)[D],
                                              double (&res)[D + 1]) {
    if constexpr (D == 1) {
        double cv[K + 1];
        load_row<K, LASTC>(c, Off(0), dd[0], cv);
        double v = 0.0, g = 0.0;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            v += dd[0].w[a] * cv[a];
            g += dd[0].dw[a] * cv[a];
        }
        res[0] = v;
        res[1] = g;
    } else if constexpr (D == 2) {
        double v = 0.0, g0 = 0.0, g1 = 0.0;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            double cv[K + 1];
            load_row<K, LASTC>(c, dd[0].off[a], dd[1], cv);
            double sB = 0.0, sD = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                sB += dd[1].w[b] * cv[b];
                sD += dd[1].dw[b] * cv[b];
            }
            v += dd[0].w[a] * sB;
            g0 += dd[0].dw[a] * sB;
            g1 += dd[0].w[a] * sD;
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
    } else {
        double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
        constexpr int kOuterUnroll = K <= INTP_OUTER_UNROLL_MAX_K ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            double tBB = 0.0, tDB = 0.0, tBD = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                double cv[K + 1];
                load_row<K, LASTC>(c, dd[0].off[0] + dd[1].off[b], dd[2], cv);
                double sB = 0.0, sD = 0.0;
#pragma unroll
                for (int e = 0; e <= K; ++e) {
                    sB += dd[2].w[e] * cv[e];
                    sD += dd[2].dw[e] * cv[e];
                }
                tBB += dd[1].w[b] * sB;
                tDB += dd[1].dw[b] * sB;
                tBD += dd[1].w[b] * sD;
            }
            v += dd[0].w[0] * tBB;
            g0 += dd[0].dw[0] * tBB;
            g1 += dd[0].w[0] * tDB;
            g2 += dd[0].w[0] * tBD;
            shift_down<K>(dd[0]);
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
        res[3] = g2;
    }
}

template <int D, int K, bool LASTC, typename Off>
__device__ __forceinline__ double contract_single(const double* __restrict__ c, DimW<K, Off> (&dd)[D]) {
    double v = 0.0;
    if constexpr (D == 1) {
        double cv[K + 1];
        load_row<K, LASTC>(c, Off(0), dd[0], cv);
#pragma unroll
        for (int a = 0; a <= K; ++a) v += dd[0].w[a] * cv[a];
    } else if constexpr (D == 2) {
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            double cv[K + 1];
            load_row<K, LASTC>(c, dd[0].off[a], dd[1], cv);
            double sum = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) sum += dd[1].w[b] * cv[b];
            v += dd[0].w[a] * sum;
        }
    } else {
        constexpr int kOuterUnroll = K <= INTP_OUTER_UNROLL_MAX_K ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            double tsum = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                double cv[K + 1];
                load_row<K, LASTC>(c, dd[0].off[0] + dd[1].off[b], dd[2], cv);
                double sum = 0.0;
#pragma unroll
                for (int e = 0; e <= K; ++e) sum += dd[2].w[e] * cv[e];
                tsum += dd[1].w[b] * sum;
            }
            v += dd[0].w[0] * tsum;
            shift_down<K>(dd[0]);
        }
    }
    return v;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Block b is remapped so that XCD x processes a contiguous range of
// chunks: with grouped points, one L2 then holds each brick patch.
__device__ __forceinline__ int64_t block_index() {
    constexpr int64_t kXcd = 8;
    int64_t b = blockIdx.x;
    const int64_t per = static_cast<int64_t>(gridDim.x) / kXcd;
    if (b < per * kXcd) b = (b % kXcd) * per + b / kXcd;
    return b;
}

template <int D, int K, bool GRAD, bool LASTC>
__device__ __forceinline__ void general_point(const EvalParams& p, int64_t pt) {
    const int64_t opt = p.perm ? static_cast<int64_t>(p.perm[pt * p.perm_stride]) : pt;
    DimW<K, int64_t> dd[D];
    bool zero = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int r = GRAD ? 1 : p.order[d];
        zero = zero || r > K;
        general_weights<K, GRAD>(p.dim[d], p.coords[d * p.cs_dim + pt * p.cs_pt], r, p.bad, dd[d]);
    }
    if constexpr (GRAD) {
        double res[D + 1];
        contract_grad<D, K, LASTC>(p.ctrl, dd, res);
#pragma unroll
        for (int s = 0; s <= D; ++s)
            if (p.slot[s] >= 0) p.out[p.slot[s] * p.os_out + opt * p.os_pt] = res[s];
    } else {
        const double v = contract_single<D, K, LASTC>(p.ctrl, dd);
        p.out[p.out_index * p.os_out + opt * p.os_pt] = zero ? 0.0 : v;
    }
}

template <int D, int K, bool GRAD, bool LASTC>
__global__ __launch_bounds__(kBlock) void eval_general_kernel(EvalParams p) {
    const int64_t pt = block_index() * kBlock + threadIdx.x;
    if (pt < p.npts) general_point<D, K, GRAD, LASTC>(p, pt);
}

template <int D, int K, bool GRAD, bool LASTC>
__device__ __forceinline__ void uniform_point(const EvalParams& p, const double (*s_poly)[(1 + 2 * K) * (K + 1) * (K + 1)],
                                              int64_t pt) {
    const int64_t opt = p.perm ? static_cast<int64_t>(p.perm[pt * p.perm_stride]) : pt;
    DimW<K, uint32_t> dd[D];
    bool zero = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int r = GRAD ? 1 : p.order[d];
        zero = zero || r > K;
        uniform_weights<K, GRAD>(p.dim[d], s_poly[d], p.coords[d * p.cs_dim + pt * p.cs_pt], r, p.bad, dd[d]);
    }
    if constexpr (GRAD) {
        double res[D + 1];
        contract_grad<D, K, LASTC>(p.ctrl, dd, res);
#pragma unroll
        for (int s = 0; s <= D; ++s)
            if (p.slot[s] >= 0) p.out[p.slot[s] * p.os_out + opt * p.os_pt] = res[s];
    } else {
        const double v = contract_single<D, K, LASTC>(p.ctrl, dd);
        p.out[p.out_index * p.os_out + opt * p.os_pt] = zero ? 0.0 : v;
    }
}

template <int D, int K, bool GRAD, bool LASTC>
__global__ __launch_bounds__(kBlock) void eval_uniform_kernel(EvalParams p) {
    constexpr int kClassMax = 1 + 2 * K;
    constexpr int kCoef = (K + 1) * (K + 1);
    __shared__ double s_poly[D][kClassMax * kCoef];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int n = p.dim[d].nclass * kCoef;
        for (int i = threadIdx.x; i < n; i += kBlock) s_poly[d][i] = p.dim[d].poly[i];
    }
    __syncthreads();
    const int64_t pt = block_index() * kBlock + threadIdx.x;
    if (pt < p.npts) uniform_point<D, K, GRAD, LASTC>(p, s_poly, pt);
}

template <void (*KF)(EvalParams)>
hipError_t launch_chunks(const EvalParams& p, hipStream_t s) {
    const int64_t blocks = (p.npts + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KF, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

template <int D, int K, bool LASTC>
hipError_t launch_variant(bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    if (uniform) {
        if (grad) return launch_chunks<eval_uniform_kernel<D, K, true, LASTC>>(p, s);
        return launch_chunks<eval_uniform_kernel<D, K, false, LASTC>>(p, s);
    }
    if (grad) return launch_chunks<eval_general_kernel<D, K, true, LASTC>>(p, s);
    return launch_chunks<eval_general_kernel<D, K, false, LASTC>>(p, s);
}

template <int D, int K>
hipError_t launch_one(bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    // contiguous stencil rows need a non-periodic last dimension (stride 1)
    if (p.dim[D - 1].periodic) return launch_variant<D, K, false>(grad, uniform, p, s);
    return launch_variant<D, K, true>(grad, uniform, p, s);
}

template <int D>
hipError_t dispatch_order(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    switch (order) {
        case 0: return launch_one<D, 0>(grad, uniform, p, s);
        case 1: return launch_one<D, 1>(grad, uniform, p, s);
        case 2: return launch_one<D, 2>(grad, uniform, p, s);
        case 3: return launch_one<D, 3>(grad, uniform, p, s);
        case 4: return launch_one<D, 4>(grad, uniform, p, s);
        case 5: return launch_one<D, 5>(grad, uniform, p, s);
        case 6: return launch_one<D, 6>(grad, uniform, p, s);
        case 7: return launch_one<D, 7>(grad, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

hipError_t dispatch(int dim, int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    switch (dim) {
        case 1: return dispatch_order<1>(order, grad, uniform, p, s);
        case 2: return dispatch_order<2>(order, grad, uniform, p, s);
        case 3: return dispatch_order<3>(order, grad, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace

hipError_t launch_eval_grad(int dim, int order, bool uniform, const EvalParams& p, hipStream_t s) {
    return dispatch(dim, order, true, uniform, p, s);
}

hipError_t launch_eval_single(int dim, int order, bool uniform, const EvalParams& p, hipStream_t s) {
    return dispatch(dim, order, false, uniform, p, s);
}

}  // namespace intp_amd
