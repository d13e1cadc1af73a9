// eval.hip -- batched evaluation of tensor-product B-splines on gfx950.
//
// One lane per query point.  Per dimension, in registers: periodic wrap
// (BSpline.hpp:130-136), knot span from the uniform-grid hint with a
// binary-search fallback (Interpolation.hpp:277-295, BSpline.hpp:125-157),
// the Cox-de Boor triangle (BSpline.hpp:83-111) with reciprocal knot gaps
// read from a per-span table, and derivative weights.  The reference
// differentiates the gathered control points (BSpline.hpp:489-520); here
// the same difference operators are applied, transposed, to the basis
// weights, so every requested output reuses one gather of the (k+1)^D
// stencil.  The stencil is contracted one axis at a time (last, contiguous
// axis first): value + all first partials cost 2(k+1)^D + 3(k+1)^(D-1) +
// ... FMAs instead of 4(k+1)^D.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kBlock = 256;

template <int K>
struct Dim1 {
    double w[K + 1];   // value weights (or the single output's weights)
    double dw[K + 1];  // first-derivative weights (gradient kernels)
    int64_t off[K + 1];
};

// Periodic wrap, span, bounds flag.  Returns mu; x is replaced by the
// wrapped coordinate.
template <int K>
__device__ __forceinline__ int locate(const DimParams& dp, double& x, int* bad) {
    if (dp.periodic) {
        double r = x - dp.lo;
        if (!(r >= 0.0 && r < dp.period)) r = fmod(r, dp.period);
        x = dp.lo + r + (x < dp.lo ? dp.period : 0.0);
    } else if (bad != nullptr && (x < dp.lo || x > dp.hi)) {
        *bad = 1;
    }
    const int last = dp.nt - K - 2;
    const double* __restrict__ t = dp.knots;
    int mu = K;
    if (dp.uniform) {
        double v = (x - dp.lo) * dp.inv_dx - dp.shift;
        v = v > 0.0 ? v : 0.0;
        mu = v >= static_cast<double>(last - K) ? last : min(last, static_cast<int>(ceil(v)) + K);
    }
    if (!(t[mu] <= x && x < t[mu + 1])) {
        if (mu > K && t[mu - 1] <= x && x < t[mu]) {
            mu -= 1;
        } else if (mu < last && t[mu + 1] <= x && x < t[mu + 2]) {
            mu += 1;
        } else {
            int lo = K + 1, hi = last + 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (x < t[mid])
                    hi = mid;
                else
                    lo = mid + 1;
            }
            mu = lo - 1;
        }
    }
    return mu;
}

// Cox-de Boor triangle on span mu.  `keep` receives level `keep_level`
// (the basis of order keep_level, right-aligned); `b` ends at level K.
template <int K>
__device__ __forceinline__ void triangle(const DimParams& dp, int mu, double x, double (&b)[K + 1],
                                         double (&keep)[K + 1], int keep_level, double (&tt)[2 * K + 1],
                                         double (&R)[(K * (K + 1)) / 2 + 1]) {
    constexpr int T = (K * (K + 1)) / 2;
    const double* __restrict__ t = dp.knots;
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

template <int K>
__device__ __forceinline__ void stencil_offsets(const DimParams& dp, int mu, int64_t (&off)[K + 1]) {
    const int64_t base = mu - K;
#pragma unroll
    for (int m = 0; m <= K; ++m) {
        int64_t id = base + m;
        if (dp.periodic && id >= dp.n) id -= dp.n;
        off[m] = id * dp.stride;
    }
}

// Gradient kernels: value + first partials from one gather.
template <int D, int K>
__global__ __launch_bounds__(kBlock) void eval_grad_kernel(EvalParams p) {
    const int64_t pt = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (pt >= p.npts) return;
    Dim1<K> dd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const DimParams& dp = p.dim[d];
        double x = p.coords[d * p.cs_dim + pt * p.cs_pt];
        const int mu = locate<K>(dp, x, p.bad);
        double tt[2 * K + 1];
        double R[(K * (K + 1)) / 2 + 1];
        double keep[K + 1];
        triangle<K>(dp, mu, x, dd[d].w, keep, K - 1, tt, R);
        if (K >= 1) {
            differentiate<K>(keep, 1, R);
#pragma unroll
            for (int m = 0; m <= K; ++m) dd[d].dw[m] = keep[m];
        } else {
            dd[d].dw[0] = 0.0;
        }
        stencil_offsets<K>(dp, mu, dd[d].off);
    }
    const double* __restrict__ c = p.ctrl;
    double res[D + 1];
    if constexpr (D == 1) {
        double v = 0.0, g = 0.0;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            const double cv = c[dd[0].off[a]];
            v += dd[0].w[a] * cv;
            g += dd[0].dw[a] * cv;
        }
        res[0] = v;
        res[1] = g;
    } else if constexpr (D == 2) {
        double v = 0.0, g0 = 0.0, g1 = 0.0;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            const double* __restrict__ row = c + dd[0].off[a];
            double sB = 0.0, sD = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                const double cv = row[dd[1].off[b]];
                sB += dd[1].w[b] * cv;
                sD += dd[1].dw[b] * cv;
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
        // high orders: keep one (k+1)^2 slab of loads live at a time
        constexpr int kOuterUnroll = K <= 3 ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            double tBB = 0.0, tDB = 0.0, tBD = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                const double* __restrict__ row = c + dd[0].off[a] + dd[1].off[b];
                double sB = 0.0, sD = 0.0;
#pragma unroll
                for (int e = 0; e <= K; ++e) {
                    const double cv = row[dd[2].off[e]];
                    sB += dd[2].w[e] * cv;
                    sD += dd[2].dw[e] * cv;
                }
                tBB += dd[1].w[b] * sB;
                tDB += dd[1].dw[b] * sB;
                tBD += dd[1].w[b] * sD;
            }
            v += dd[0].w[a] * tBB;
            g0 += dd[0].dw[a] * tBB;
            g1 += dd[0].w[a] * tDB;
            g2 += dd[0].w[a] * tBD;
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
        res[3] = g2;
    }
#pragma unroll
    for (int s = 0; s <= D; ++s)
        if (p.slot[s] >= 0) p.out[p.slot[s] * p.os_out + pt * p.os_pt] = res[s];
}

// Single arbitrary mixed partial (orders may exceed K: the result is 0,
// BSpline.hpp:404-407).
template <int D, int K>
__global__ __launch_bounds__(kBlock) void eval_single_kernel(EvalParams p) {
    const int64_t pt = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (pt >= p.npts) return;
    Dim1<K> dd[D];
    bool zero = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const DimParams& dp = p.dim[d];
        const int r = p.order[d];
        zero = zero || r > K;
        double x = p.coords[d * p.cs_dim + pt * p.cs_pt];
        const int mu = locate<K>(dp, x, p.bad);
        double tt[2 * K + 1];
        double R[(K * (K + 1)) / 2 + 1];
        double keep[K + 1];
        double b[K + 1];
        triangle<K>(dp, mu, x, b, keep, K - r, tt, R);
        if (r == 0) {
#pragma unroll
            for (int m = 0; m <= K; ++m) keep[m] = b[m];
        } else if (r <= K) {
            differentiate<K>(keep, r, R);
        }
#pragma unroll
        for (int m = 0; m <= K; ++m) dd[d].w[m] = keep[m];
        stencil_offsets<K>(dp, mu, dd[d].off);
    }
    const double* __restrict__ c = p.ctrl;
    double v = 0.0;
    if constexpr (D == 1) {
#pragma unroll
        for (int a = 0; a <= K; ++a) v += dd[0].w[a] * c[dd[0].off[a]];
    } else if constexpr (D == 2) {
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            const double* __restrict__ row = c + dd[0].off[a];
            double s = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) s += dd[1].w[b] * row[dd[1].off[b]];
            v += dd[0].w[a] * s;
        }
    } else {
        constexpr int kOuterUnroll = K <= 3 ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            double tsum = 0.0;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                const double* __restrict__ row = c + dd[0].off[a] + dd[1].off[b];
                double s = 0.0;
#pragma unroll
                for (int e = 0; e <= K; ++e) s += dd[2].w[e] * row[dd[2].off[e]];
                tsum += dd[1].w[b] * s;
            }
            v += dd[0].w[a] * tsum;
        }
    }
    p.out[p.out_index * p.os_out + pt * p.os_pt] = zero ? 0.0 : v;
}

template <int D, int K>
hipError_t launch_one(bool grad, const EvalParams& p, hipStream_t s) {
    const int64_t blocks = (p.npts + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (grad)
        hipLaunchKernelGGL((eval_grad_kernel<D, K>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    else
        hipLaunchKernelGGL((eval_single_kernel<D, K>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

template <int D>
hipError_t dispatch_order(int order, bool grad, const EvalParams& p, hipStream_t s) {
    switch (order) {
        case 0: return launch_one<D, 0>(grad, p, s);
        case 1: return launch_one<D, 1>(grad, p, s);
        case 2: return launch_one<D, 2>(grad, p, s);
        case 3: return launch_one<D, 3>(grad, p, s);
        case 4: return launch_one<D, 4>(grad, p, s);
        case 5: return launch_one<D, 5>(grad, p, s);
        case 6: return launch_one<D, 6>(grad, p, s);
        case 7: return launch_one<D, 7>(grad, p, s);
        default: return hipErrorNotSupported;
    }
}

hipError_t dispatch(int dim, int order, bool grad, const EvalParams& p, hipStream_t s) {
    switch (dim) {
        case 1: return dispatch_order<1>(order, grad, p, s);
        case 2: return dispatch_order<2>(order, grad, p, s);
        case 3: return dispatch_order<3>(order, grad, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace

hipError_t launch_eval_grad(int dim, int order, const EvalParams& p, hipStream_t s) {
    return dispatch(dim, order, true, p, s);
}

hipError_t launch_eval_single(int dim, int order, const EvalParams& p, hipStream_t s) {
    return dispatch(dim, order, false, p, s);
}

}  // namespace intp_amd
