// eval_impl.hpp -- batched evaluation of tensor-product B-splines on
// gfx950 (kernel templates; eval_f64_d*.hip / eval_f32_d*.hip instantiate
// them per table type and dimension).
//
// One lane per query point.  Per dimension, in registers: periodic wrap
// (BSpline.hpp:130-136) and the knot span from the uniform-grid hint with a
// binary-search fallback (Interpolation.hpp:277-295, BSpline.hpp:125-157),
// then the basis weights:
//  * uniform grids (fast path): knots are recomputed from their closed form
//    with the host's exact operation order (no FMA), so the span test is
//    the reference's, with no memory traffic; the basis comes from the
//    span's polynomial coefficients (one class for all interior spans,
//    one per clamped end span), read from the device table;
//  * other grids: knots and reciprocal knot gaps are read from HBM and the
//    Cox-de Boor triangle (BSpline.hpp:83-111) runs in registers.
// Derivatives: the reference differentiates the gathered control points
// (BSpline.hpp:489-520); here the same difference operators act,
// transposed, on the basis weights (or the basis polynomials are
// differentiated), so all requested outputs share one gather of the
// (k+1)^D stencil.  The stencil is contracted one axis at a time, last
// (contiguous) axis first.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "span.hpp"

namespace intp_amd {

namespace {

constexpr int kEvalBlock = 256;

#ifndef INTP_OUTER_UNROLL_MAX_K
#define INTP_OUTER_UNROLL_MAX_K 2  // fully unroll the outer stencil loop up to this order
#endif

template <int K, typename Off, typename W>
struct DimW {
    W w[K + 1];      // value weights (or the single output's weights)
    W dw[K + 1];     // first-derivative weights (gradient kernels); on a
                     // gradient table, the weights of the differences
    Off off[K + 1];  // element offsets of the K+1 stencil planes
};

// Weight and result type of a table element type: the element type
// itself, or float for the cells of a refined build's gradient table.
template <typename T>
struct TableWeight {
    using type = T;
};
template <>
struct TableWeight<Cell1> {
    using type = float;
};
template <>
struct TableWeight<Cell3> {
    using type = float;
};
template <typename T>
using weight_t = typename TableWeight<T>::type;
template <typename T>
constexpr bool is_cells() {
    return !std::is_same<T, weight_t<T>>::value;
}

// Weights of the backward differences from the first-derivative weights
// dw of one dimension: sum_m dw[m] c[m] = sum_m g[m] (c[m] - c[m-1]) with
// g[m] = -(dw[0] + ... + dw[m-1]) (summation by parts; the dw sum to 0),
// so g[0] = 0 and g[m] = k B_{k-1}[m] / (t[mu+m] - t[mu-k+m]) >= 0, the
// scaled order-(k-1) basis of the derivative identity.  Summed in FP64.
template <int K, typename W>
__device__ __forceinline__ void difference_weights(const double (&dw)[K + 1], W (&g)[K + 1]) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m <= K; ++m) {
        g[m] = static_cast<W>(-s);
        s += dw[m];
    }
}

// ---------------------------------------------------------------------------
// span search
// ---------------------------------------------------------------------------

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
template <int K, bool GRAD, typename W, bool DIFF = false>
__device__ __forceinline__ void general_weights(const DimParams& dp, double x, int r, int* bad,
                                                DimW<K, int64_t, W>& dd) {
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
        for (int m = 0; m <= K; ++m) dd.w[m] = static_cast<W>(b[m]);
        if (K >= 1) differentiate<K>(keep, 1, R);
#pragma unroll
        for (int m = 0; m <= K; ++m) keep[m] = K >= 1 ? keep[m] : 0.0;
        if constexpr (DIFF) {
            difference_weights<K, W>(keep, dd.dw);
        } else {
#pragma unroll
            for (int m = 0; m <= K; ++m) dd.dw[m] = static_cast<W>(keep[m]);
        }
    } else {
        if (r == 0) {
#pragma unroll
            for (int m = 0; m <= K; ++m) keep[m] = b[m];
        } else if (r <= K) {
            differentiate<K>(keep, r, R);
        }
#pragma unroll
        for (int m = 0; m <= K; ++m) dd.w[m] = static_cast<W>(keep[m]);
    }
    stencil_offsets<K>(dp, mu, dd.off);
}

// ---------------------------------------------------------------------------
// basis weights, uniform grids
// ---------------------------------------------------------------------------

template <int K>
__device__ __forceinline__ double uknot_sel(const DimParams& dp, int i) {
    return uknot_sel(dp, K, i);
}

template <int K>
__device__ __forceinline__ int uniform_span(const DimParams& dp, double x) {
    return uniform_span(dp, K, x);
}

// Weights from the span's basis polynomials P (row j: coefficients of
// basis j in u = (x - t_mu) / dx, ascending powers); PP is a pointer to
// the coefficients in a given address space.
template <int K, bool GRAD, typename W, bool DIFF, typename PP>
__device__ __forceinline__ void poly_weights(PP P, double u, double inv_dx, int r, DimW<K, uint32_t, W>& dd) {
    constexpr int NW = K + 1;
    if constexpr (GRAD) {
        double sum = 0.0;  // DIFF: running sum of the derivative weights (difference_weights)
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            double v = P[j * NW + K];
            double g = static_cast<double>(K) * P[j * NW + K];
#pragma unroll
            for (int q = K - 1; q >= 0; --q) {
                v = v * u + P[j * NW + q];
                if (q >= 1) g = g * u + static_cast<double>(q) * P[j * NW + q];
            }
            dd.w[j] = static_cast<W>(v);
            const double dw = K >= 1 ? g * inv_dx : 0.0;
            if constexpr (DIFF) {
                dd.dw[j] = static_cast<W>(-sum);
                sum += dw;
            } else {
                dd.dw[j] = static_cast<W>(dw);
            }
        }
    } else if (r == 0) {
        // the value: Horner
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            double v = P[j * NW + K];
#pragma unroll
            for (int q = K - 1; q >= 0; --q) v = v * u + P[j * NW + q];
            dd.w[j] = static_cast<W>(v);
        }
    } else {
        // r-th derivative: sum_q q!/(q-r)! P_q u^(q-r), times dx^-r
        double scale = 1.0;
        for (int i = 0; i < r; ++i) scale *= inv_dx;
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            double v = 0.0;
#pragma unroll
            for (int q = K; q >= 0; --q) {
                if (q >= r) {
                    double ff = 1.0;
#pragma unroll
                    for (int t = 0; t < K; ++t)
                        if (t < r) ff *= static_cast<double>(q - t);
                    v = v * u + ff * P[j * NW + q];
                }
            }
            dd.w[j] = static_cast<W>(r > K ? 0.0 : v * scale);
        }
    }
}

// Returns the span; fills the weights and the HBM offsets of the stencil.
// The basis polynomials of the span's class come from dp.poly: with scalar
// loads when every lane of the wave is on an interior span (class 0, the
// common case), otherwise with vector loads of each lane's class (a few KB,
// cached).  No LDS copy: staging the classes in every workgroup cost more
// than it saved (up to 9.5 KB per workgroup at order 5).
template <int K, bool GRAD, typename W, bool DIFF = false>
__device__ __forceinline__ int uniform_weights(const DimParams& dp, double x, int r, int* bad,
                                               DimW<K, uint32_t, W>& dd) {
    wrap_or_check(dp, x, bad);
    const int last = dp.nt - K - 2;
    const int mu = uniform_span<K>(dp, x);
    const int cls = (mu - K < dp.nleft) ? 1 + (mu - K) : ((last - mu < dp.nright) ? 1 + dp.nleft + (last - mu) : 0);
    const double u = (x - uknot_sel<K>(dp, mu)) * dp.inv_dx;
    // explicit address spaces: as one generic pointer both branches would
    // read through flat loads (and the uniform case without scalar loads)
    if (__all(cls == 0))
        poly_weights<K, GRAD, W, DIFF>((const __attribute__((address_space(4))) double*)dp.poly, u, dp.inv_dx, r, dd);
    else
        poly_weights<K, GRAD, W, DIFF>(
            (const __attribute__((address_space(1))) double*)(dp.poly + cls * (K + 1) * (K + 1)), u, dp.inv_dx, r, dd);
    stencil_offsets<K>(dp, mu, dd.off);
    return mu;
}

// ---------------------------------------------------------------------------
// stencil contraction
// ---------------------------------------------------------------------------

// One stencil row along the last dimension.  LASTC: the last dimension is
// not periodic, so the row is K+1 consecutive elements from one address
// (one address computation, immediate offsets).
template <int K, bool LASTC, typename T, typename Off, typename W>
__device__ __forceinline__ void load_row(const T* __restrict__ c, Off rowoff, const DimW<K, Off, W>& dl,
                                         T (&cv)[K + 1]) {
    if constexpr (LASTC) {
        const T* __restrict__ r = c + (rowoff + dl.off[0]);
#pragma unroll
        for (int e = 0; e <= K; ++e) cv[e] = r[e];
    } else {
        const T* __restrict__ r = c + rowoff;
#pragma unroll
        for (int e = 0; e <= K; ++e) cv[e] = r[dl.off[e]];
    }
}

// Shifts the weights/offsets of the first dimension down by one, so a
// rolled loop over its K+1 planes only ever reads element 0 (registers,
// never scratch).
template <int K, typename Off, typename W>
__device__ __forceinline__ void shift_down(DimW<K, Off, W>& d) {
#pragma unroll
    for (int m = 0; m < K; ++m) {
        d.w[m] = d.w[m + 1];
        d.dw[m] = d.dw[m + 1];
        d.off[m] = d.off[m + 1];
    }
}

// 4 to 6 dimensions: the outer D-1 dimensions run as one rolled loop
// over their (K+1)^(D-1) index combinations (a counter with the last outer
// dimension fastest), each contracting one stencil row along the last
// dimension.  The weights are indexed at run time, so these kernels keep
// them in scratch memory.  GRAD: res = value and the D first partials
// (dw along one dimension, w along the others); otherwise res[0] only.
template <int D, int K, bool GRAD, bool LASTC, typename T, typename Off>
__device__ void contract_nd(const T* __restrict__ c, DimW<K, Off, T> (&dd)[D], T (&res)[D + 1]) {
    int a[D - 1];
#pragma unroll
    for (int d = 0; d < D - 1; ++d) a[d] = 0;
#pragma unroll
    for (int s = 0; s <= D; ++s) res[s] = T(0);
    for (;;) {
        Off rowoff = 0;
        T wv = T(1);
#pragma unroll
        for (int d = 0; d < D - 1; ++d) {
            rowoff += dd[d].off[a[d]];
            wv *= dd[d].w[a[d]];
        }
        T cv[K + 1];
        load_row<K, LASTC, T>(c, rowoff, dd[D - 1], cv);
        T sB = T(0), sD = T(0);
#pragma unroll
        for (int e = 0; e <= K; ++e) {
            sB += dd[D - 1].w[e] * cv[e];
            if constexpr (GRAD) sD += dd[D - 1].dw[e] * cv[e];
        }
        res[0] += wv * sB;
        if constexpr (GRAD) {
#pragma unroll
            for (int d = 0; d < D - 1; ++d) {
                T wd = T(1);
#pragma unroll
                for (int e = 0; e < D - 1; ++e) wd *= e == d ? dd[e].dw[a[e]] : dd[e].w[a[e]];
                res[d + 1] += wd * sB;
            }
            res[D] += wv * sD;
        }
        int d = D - 2;
        for (; d >= 0; --d) {
            if (++a[d] <= K) break;
            a[d] = 0;
        }
        if (d < 0) break;
    }
}

template <int D, int K, bool LASTC, typename T, typename Off>
__device__ __forceinline__ void contract_grad(const T* __restrict__ c, DimW<K, Off, T> (&dd)[D], T (&res)[D + 1]) {
    if constexpr (D == 1) {
        T cv[K + 1];
        load_row<K, LASTC, T>(c, Off(0), dd[0], cv);
        T v = T(0), g = T(0);
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            v += dd[0].w[a] * cv[a];
            g += dd[0].dw[a] * cv[a];
        }
        res[0] = v;
        res[1] = g;
    } else if constexpr (D == 2) {
        T v = T(0), g0 = T(0), g1 = T(0);
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            T cv[K + 1];
            load_row<K, LASTC, T>(c, dd[0].off[a], dd[1], cv);
            T sB = T(0), sD = T(0);
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
    } else if constexpr (D >= 4) {
        contract_nd<D, K, true, LASTC>(c, dd, res);
    } else {
        T v = T(0), g0 = T(0), g1 = T(0), g2 = T(0);
        constexpr int kOuterUnroll = K <= INTP_OUTER_UNROLL_MAX_K ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            T tBB = T(0), tDB = T(0), tBD = T(0);
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                T cv[K + 1];
                load_row<K, LASTC, T>(c, dd[0].off[0] + dd[1].off[b], dd[2], cv);
                T sB = T(0), sD = T(0);
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
            shift_down<K, Off, T>(dd[0]);
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
        res[3] = g2;
    }
}

template <int D, int K, bool LASTC, typename T, typename Off>
__device__ __forceinline__ T contract_single(const T* __restrict__ c, DimW<K, Off, T> (&dd)[D]) {
    T v = T(0);
    if constexpr (D == 1) {
        T cv[K + 1];
        load_row<K, LASTC, T>(c, Off(0), dd[0], cv);
#pragma unroll
        for (int a = 0; a <= K; ++a) v += dd[0].w[a] * cv[a];
    } else if constexpr (D == 2) {
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            T cv[K + 1];
            load_row<K, LASTC, T>(c, dd[0].off[a], dd[1], cv);
            T sum = T(0);
#pragma unroll
            for (int b = 0; b <= K; ++b) sum += dd[1].w[b] * cv[b];
            v += dd[0].w[a] * sum;
        }
    } else if constexpr (D >= 4) {
        T res[D + 1];
        contract_nd<D, K, false, LASTC>(c, dd, res);
        v = res[0];
    } else {
        constexpr int kOuterUnroll = K <= INTP_OUTER_UNROLL_MAX_K ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            T tsum = T(0);
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                T cv[K + 1];
                load_row<K, LASTC, T>(c, dd[0].off[0] + dd[1].off[b], dd[2], cv);
                T sum = T(0);
#pragma unroll
                for (int e = 0; e <= K; ++e) sum += dd[2].w[e] * cv[e];
                tsum += dd[1].w[b] * sum;
            }
            v += dd[0].w[0] * tsum;
            shift_down<K, Off, T>(dd[0]);
        }
    }
    return v;
}

// Value and first partials on a gradient table (cells {c, e_0, .., e_{D-1}}):
// the value contracts c with the value weights w, partial d contracts e_d
// with the difference weights (dw) along d and w along the others.  Same
// axis order as contract_grad.
template <int D, int K, bool LASTC, typename T, typename Off>
__device__ __forceinline__ void contract_cells(const T* __restrict__ c, DimW<K, Off, float> (&dd)[D],
                                               float (&res)[D + 1]) {
    if constexpr (D == 1) {
        T cv[K + 1];
        load_row<K, LASTC, T>(c, Off(0), dd[0], cv);
        float v = 0.f, g = 0.f;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            v += dd[0].w[a] * cv[a].c;
            g += dd[0].dw[a] * cv[a].e0;
        }
        res[0] = v;
        res[1] = g;
    } else if constexpr (D == 2) {
        float v = 0.f, g0 = 0.f, g1 = 0.f;
#pragma unroll
        for (int a = 0; a <= K; ++a) {
            T cv[K + 1];
            load_row<K, LASTC, T>(c, dd[0].off[a], dd[1], cv);
            float sC = 0.f, s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                sC += dd[1].w[b] * cv[b].c;
                s0 += dd[1].w[b] * cv[b].e0;
                s1 += dd[1].dw[b] * cv[b].e1;
            }
            v += dd[0].w[a] * sC;
            g0 += dd[0].dw[a] * s0;
            g1 += dd[0].w[a] * s1;
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
    } else {
        float v = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
        constexpr int kOuterUnroll = K <= INTP_OUTER_UNROLL_MAX_K ? K + 1 : 1;
#pragma unroll kOuterUnroll
        for (int a = 0; a <= K; ++a) {
            float tC = 0.f, t0 = 0.f, t1 = 0.f, t2 = 0.f;
#pragma unroll
            for (int b = 0; b <= K; ++b) {
                T cv[K + 1];
                load_row<K, LASTC, T>(c, dd[0].off[0] + dd[1].off[b], dd[2], cv);
                float sC = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
                for (int e = 0; e <= K; ++e) {
                    sC += dd[2].w[e] * cv[e].c;
                    s0 += dd[2].w[e] * cv[e].e0;
                    s1 += dd[2].w[e] * cv[e].e1;
                    s2 += dd[2].dw[e] * cv[e].e2;
                }
                tC += dd[1].w[b] * sC;
                t0 += dd[1].w[b] * s0;
                t1 += dd[1].dw[b] * s1;
                t2 += dd[1].w[b] * s2;
            }
            v += dd[0].w[0] * tC;
            g0 += dd[0].dw[0] * t0;
            g1 += dd[0].w[0] * t1;
            g2 += dd[0].w[0] * t2;
            shift_down<K, Off, float>(dd[0]);
        }
        res[0] = v;
        res[1] = g0;
        res[2] = g1;
        res[3] = g2;
    }
}

// The gradient kernels' contraction for a table of element type T.
template <int D, int K, bool LASTC, typename T, typename Off>
__device__ __forceinline__ void contract_outputs(const T* __restrict__ c, DimW<K, Off, weight_t<T>> (&dd)[D],
                                                 weight_t<T> (&res)[D + 1]) {
    if constexpr (is_cells<T>())
        contract_cells<D, K, LASTC, T, Off>(c, dd, res);
    else
        contract_grad<D, K, LASTC, T, Off>(c, dd, res);
}

// Stores the value and first partials of one point.
// One output value (FP64, or FP32 into out32).
__device__ __forceinline__ void store_out(const EvalParams& p, int64_t at, double v) {
    if (p.out32)
        p.out32[at] = static_cast<float>(v);
    else
        p.out[at] = v;
}

template <int D, typename T>
__device__ __forceinline__ void store_grad(const EvalParams& p, int64_t opt, const T (&res)[D + 1]) {
    if (p.out_packed && !p.out32) {
        double* o = p.out + opt * p.os_pt;
#pragma unroll
        for (int s = 0; s + 1 <= D; s += 2) *reinterpret_cast<double2*>(o + s) = make_double2(res[s], res[s + 1]);
        if constexpr ((D + 1) % 2 == 1) o[D] = res[D];
    } else {
#pragma unroll
        for (int s = 0; s <= D; ++s)
            if (p.slot[s] >= 0) store_out(p, p.slot[s] * p.os_out + opt * p.os_pt, res[s]);
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Block b is remapped so that XCD x processes a contiguous range of
// chunks: with grouped points, one L2 then holds each brick patch.
__device__ __forceinline__ int64_t block_index() { return xcd_chunk_index(); }

template <int D, int K, bool GRAD, bool LASTC, typename T>
__device__ __forceinline__ void general_point(const EvalParams& p, int64_t pt) {
    using W = weight_t<T>;
    const int64_t opt = p.perm ? static_cast<int64_t>(p.perm[pt * p.perm_stride]) : pt;
    const T* __restrict__ ctrl = static_cast<const T*>(p.ctrl);
    DimW<K, int64_t, W> dd[D];
    bool zero = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int r = GRAD ? 1 : p.order[d];
        zero = zero || r > K;
        general_weights<K, GRAD, W, is_cells<T>()>(p.dim[d], p.coords[d * p.cs_dim + pt * p.cs_pt], r, p.bad,
                                                   dd[d]);
    }
    if constexpr (GRAD) {
        W res[D + 1];
        contract_outputs<D, K, LASTC>(ctrl, dd, res);
#pragma unroll
        for (int s = 0; s <= D; ++s)
            if (p.slot[s] >= 0) store_out(p, p.slot[s] * p.os_out + opt * p.os_pt, res[s]);
    } else {
        const T v = contract_single<D, K, LASTC>(ctrl, dd);
        store_out(p, p.out_index * p.os_out + opt * p.os_pt, zero ? 0.0 : static_cast<double>(v));
    }
}

template <int D, int K, bool GRAD, bool LASTC, typename T>
__global__ __launch_bounds__(kEvalBlock) void eval_general_kernel(EvalParams p) {
    const int64_t pt = block_index() * kEvalBlock + threadIdx.x;
    if (pt < p.npts) general_point<D, K, GRAD, LASTC, T>(p, pt);
}

template <int D, int K, bool GRAD, bool LASTC, typename T>
__device__ __forceinline__ void uniform_point(const EvalParams& p, int64_t pt) {
    using W = weight_t<T>;
    const int64_t opt = p.perm ? static_cast<int64_t>(p.perm[pt * p.perm_stride]) : pt;
    const T* __restrict__ ctrl = static_cast<const T*>(p.ctrl);
    DimW<K, uint32_t, W> dd[D];
    bool zero = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int r = GRAD ? 1 : p.order[d];
        zero = zero || r > K;
        uniform_weights<K, GRAD, W, is_cells<T>()>(p.dim[d], p.coords[d * p.cs_dim + pt * p.cs_pt], r, p.bad,
                                                   dd[d]);
    }
    if constexpr (GRAD) {
        W res[D + 1];
        contract_outputs<D, K, LASTC>(ctrl, dd, res);
        store_grad<D>(p, opt, res);
    } else {
        const T v = contract_single<D, K, LASTC>(ctrl, dd);
        store_out(p, p.out_index * p.os_out + opt * p.os_pt, zero ? 0.0 : static_cast<double>(v));
    }
}

template <int D, int K, bool GRAD, bool LASTC, typename T>
__global__ __launch_bounds__(kEvalBlock) void eval_uniform_kernel(EvalParams p) {
    const int64_t pt = block_index() * kEvalBlock + threadIdx.x;
    if (pt < p.npts) uniform_point<D, K, GRAD, LASTC, T>(p, pt);
}

// ---- grouped evaluation with the brick patch in LDS ---------------------

// One grouped record as two 16-byte loads (the patch kernels read the
// records of the grouping passes: p.coords points at them).
__device__ __forceinline__ BinRecord load_record(const BinRecord* src) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* s = reinterpret_cast<const u32x4*>(src);
    union {
        u32x4 v[2];
        BinRecord r;
    } u;
    u.v[0] = s[0];
    u.v[1] = s[1];
    return u.r;
}

template <typename T>
constexpr int patch_max() {
    return kPatchBytes / static_cast<int>(sizeof(T));
}

// One point of a grouped chunk: if its stencil lies in the staged patch
// (its brick is the chunk's), the stencil is read from LDS, otherwise from
// HBM.  Same weights and contraction order either way.
template <int D, int K, bool GRAD, bool LASTC, typename T>
__device__ __forceinline__ void patch_point(const EvalParams& p, const T* s_patch, const int (&bstart)[D], const int (&bsize)[D],
                                            const uint32_t (&lstride)[D], const BinRecord& rec) {
    using W = weight_t<T>;
    const int64_t opt = static_cast<int64_t>(rec.idx);
    const T* __restrict__ ctrl = static_cast<const T*>(p.ctrl);
    DimW<K, uint32_t, W> dd[D];
    bool zero = false, inside = true;
    int local[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int r = GRAD ? 1 : p.order[d];
        zero = zero || r > K;
        const int mu = uniform_weights<K, GRAD, W, is_cells<T>()>(p.dim[d], rec.x[d], r, p.bad, dd[d]);
        const int n = static_cast<int>(p.dim[d].n);
        int st = mu - K;  // first stencil point; periodic even orders: may be n (wraps to 0)
        if (st >= n) st -= n;
        int l = st - bstart[d];
        if (p.dim[d].periodic && l < 0) l += n;
        local[d] = l;
        inside = inside && l >= 0 && l < bsize[d];
    }
    W res[D + 1];
    if (inside) {
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int m = 0; m <= K; ++m) dd[d].off[m] = static_cast<uint32_t>(local[d] + m) * lstride[d];
        if constexpr (GRAD)
            contract_outputs<D, K, true>(s_patch, dd, res);
        else
            res[0] = contract_single<D, K, true>(s_patch, dd);
    } else {
        if constexpr (GRAD)
            contract_outputs<D, K, LASTC>(ctrl, dd, res);
        else
            res[0] = contract_single<D, K, LASTC>(ctrl, dd);
    }
    if constexpr (GRAD) {
        store_grad<D>(p, opt, res);
    } else {
        store_out(p, p.out_index * p.os_out + opt * p.os_pt, zero ? 0.0 : static_cast<double>(res[0]));
    }
}

// One workgroup per 256 grouped points: stages the brick of its middle
// point, (B + K)^D control points (wrapped on periodic dimensions, zero
// past the end of clamped ones, where no stencil reaches), then evaluates.
template <int D, int K, bool GRAD, bool LASTC, typename T>
__global__ __launch_bounds__(kEvalBlock) void eval_patch_kernel(EvalParams p) {
    constexpr int kPatchMax = patch_max<T>();
    const T* __restrict__ ctrl = static_cast<const T*>(p.ctrl);
    __shared__ T s_patch[kPatchMax];
    const int64_t first = block_index() * kEvalBlock;
    const int64_t pivot = min(first + kEvalBlock / 2, p.npts - 1);
    const int64_t pt = first + threadIdx.x;
    uint32_t rem = p.key[pivot * p.perm_stride];
    int bstart[D], bsize[D], psize[D];
    uint32_t lstride[D];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const uint32_t nb = static_cast<uint32_t>(p.bricks.nbrick[d]);
        const uint32_t b = rem % nb;
        rem /= nb;
        bstart[d] = static_cast<int>(b << p.bricks.brick_shift[d]);
        bsize[d] = 1 << p.bricks.brick_shift[d];
        psize[d] = bsize[d] + K;
    }
    uint32_t vol = 1;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        lstride[d] = vol;
        vol *= static_cast<uint32_t>(psize[d]);
    }
    if (vol > static_cast<uint32_t>(kPatchMax) || psize[D - 1] > (D == 3 ? 16 : 64) || (D == 3 && psize[D == 3 ? 1 : 0] > 16) ||
        psize[0] > (D == 3 ? 16 : 64)) {
        // not staged: every point reads HBM
        vol = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) bsize[d] = 0;
    }
    // Each thread owns one position of the inner dimensions (16 x 16 for
    // D = 3, 64 for D = 2, so their patch extents must fit) and walks
    // dimension 0 with a running index: no divisions in the loop.
    if (vol > 0) {
        constexpr int kInner = D == 3 ? 16 : 64;     // threads along the last dimension
        constexpr int kGroups = kEvalBlock / (D == 3 ? 256 : 64);  // D = 2: 4 groups split dimension 0
        const int t = threadIdx.x;
        const int i_last = t % kInner;
        const int i_mid = D == 3 ? t / kInner : 0;
        const int group = D == 3 ? 0 : t / kInner;
        bool ok = i_last < psize[D - 1] && (D == 2 || i_mid < psize[D == 3 ? 1 : 0]);
        // wrapped element of a dimension, or -1 past the end of a clamped one
        auto element = [&](int d, int i) -> int64_t {
            int64_t g = bstart[d] + static_cast<int64_t>(i);
            const int64_t n = p.dim[d].n;
            if (g < n) return g;
            if (!p.dim[d].periodic) return -1;
            while (g >= n) g -= n;
            return g;
        };
        uint32_t inner_off = 0;
        uint32_t lds = 0;
        if (ok) {
            const int64_t gl = element(D - 1, i_last);
            ok = gl >= 0;
            inner_off = static_cast<uint32_t>(gl);
            lds = static_cast<uint32_t>(i_last);
            if constexpr (D == 3) {
                const int64_t gm = element(1, i_mid);
                ok = ok && gm >= 0;
                inner_off += static_cast<uint32_t>(gm * p.dim[1].stride);
                lds += static_cast<uint32_t>(i_mid) * lstride[1];
            }
        }
        if (ok) {
            // all loads first, then all LDS stores, so the loads overlap
            constexpr int kRows = D == 3 ? 16 : 16;  // >= psize[0] / kGroups
            const int n0 = static_cast<int>(p.dim[0].n);
            const bool per0 = p.dim[0].periodic != 0;
            const uint32_t stride0 = static_cast<uint32_t>(p.dim[0].stride);
            T v[kRows];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const int i0 = group + r * kGroups;
                v[r] = T{};
                if (i0 < psize[0]) {
                    int g0 = bstart[0] + i0;
                    if (g0 >= n0) g0 = per0 ? g0 % n0 : -1;
                    if (g0 >= 0) v[r] = ctrl[inner_off + static_cast<uint32_t>(g0) * stride0];
                }
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const int i0 = group + r * kGroups;
                if (i0 < psize[0]) s_patch[lds + static_cast<uint32_t>(i0) * lstride[0]] = v[r];
            }
        }
        // positions past the end of a clamped dimension stay unwritten: no
        // stencil of a point inside the brick reaches them
    }
    __syncthreads();
    // the thread's grouped record (coordinates, caller position); loaded
    // after the barrier (issued before the staging, it measured slower)
    if (pt < p.npts)
        patch_point<D, K, GRAD, LASTC, T>(p, s_patch, bstart, bsize, lstride,
                                          load_record(reinterpret_cast<const BinRecord*>(p.coords) + pt));
}

// One workgroup per brick item (one-pass grouping): the item's records
// all belong to one brick, whose (B_d + K) patch of control points is
// staged once in dynamic LDS (wrapped on periodic dimensions, zero past
// the end of clamped ones, where no stencil of the brick reaches); each
// thread's kBrickPointsPerThread records are loaded before the staging so
// that both sets of loads are in flight together.
// Waves per SIMD the brick kernel is built for: 4 (at most 128 VGPRs)
// up to order 3, else 2.
template <int D, int K>
constexpr int brick_waves() {
    return K <= 3 ? 4 : 2;
}

template <int D, int K, bool GRAD, bool LASTC, typename T>
__global__ __launch_bounds__(kBrickThreads, (brick_waves<D, K>())) void eval_brick_kernel(EvalParams p) {
    extern __shared__ __align__(16) unsigned char brick_smem[];
    T* s_patch = reinterpret_cast<T*>(brick_smem);
    const int64_t item_id = blockIdx.x;
    if (item_id >= static_cast<int64_t>(*p.item_count)) return;
    const BrickItem it = p.items[item_id];
    const BinRecord* __restrict__ recs = reinterpret_cast<const BinRecord*>(p.coords);
    // the thread's first record is in flight during the staging; each
    // later one during the previous point's evaluation
    BinRecord next{};
    if (threadIdx.x < it.count) next = load_record(recs + it.start + threadIdx.x);
    // brick of the item
    int bstart[D], bsize[D], psize[D];
    uint32_t lstride[D];
    uint32_t rem = it.bin;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const uint32_t nb = static_cast<uint32_t>(p.bricks.nbrick[d]);
        const uint32_t b = rem % nb;
        rem /= nb;
        bstart[d] = static_cast<int>(b << p.bricks.brick_shift[d]);
        bsize[d] = 1 << p.bricks.brick_shift[d];
        psize[d] = bsize[d] + K;
    }
    uint32_t vol = 1;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        lstride[d] = vol;
        vol *= static_cast<uint32_t>(psize[d]);
    }
    const T* __restrict__ ctrl = static_cast<const T*>(p.ctrl);
    if (vol <= static_cast<uint32_t>(p.patch_volume) && psize[D - 1] <= kBrickThreads) {
        // rows along the last dimension: thread t owns position t % P of a
        // row and walks rows t / P, t / P + R, ... (R = threads / P)
        const int P = psize[D - 1];
        const int R = kBrickThreads / P;
        const int i_last = threadIdx.x % P;
        int row = threadIdx.x / P;
        const int nrows = static_cast<int>(vol) / P;
        int64_t gl = bstart[D - 1] + i_last;
        const int64_t n_last = p.dim[D - 1].n;
        bool ok_last = row < R;
        if (gl >= n_last) {
            if (p.dim[D - 1].periodic)
                while (gl >= n_last) gl -= n_last;
            else
                ok_last = false;
        }
        if (ok_last) {
            for (; row < nrows; row += R) {
                // row -> outer indices (C order over dimensions 0..D-2)
                int r = row;
                int64_t goff = gl;
                bool ok = true;
#pragma unroll
                for (int d = D - 2; d >= 0; --d) {
                    const int i = r % psize[d];
                    r /= psize[d];
                    int64_t g = bstart[d] + i;
                    if (g >= p.dim[d].n) {
                        if (p.dim[d].periodic)
                            while (g >= p.dim[d].n) g -= p.dim[d].n;
                        else
                            ok = false;
                    }
                    goff += g * p.dim[d].stride;
                }
                if (ok) s_patch[static_cast<uint32_t>(row) * static_cast<uint32_t>(P) + i_last] = ctrl[goff];
            }
        }
    } else {
        // no room: every point reads HBM
#pragma unroll
        for (int d = 0; d < D; ++d) bsize[d] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBrickPointsPerThread; ++j) {
        const uint32_t q = j * kBrickThreads + threadIdx.x;
        const BinRecord cur = next;
        if (j + 1 < kBrickPointsPerThread && q + kBrickThreads < it.count)
            next = load_record(recs + it.start + q + kBrickThreads);
        if (q < it.count) patch_point<D, K, GRAD, LASTC, T>(p, s_patch, bstart, bsize, lstride, cur);
    }
}

template <void (*KF)(EvalParams)>
hipError_t launch_chunks(const EvalParams& p, hipStream_t s) {
    const int64_t blocks = (p.npts + kEvalBlock - 1) / kEvalBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KF, dim3(static_cast<unsigned>(blocks)), dim3(kEvalBlock), 0, s, p);
    return hipGetLastError();
}

template <void (*KF)(EvalParams)>
hipError_t launch_items(const EvalParams& p, size_t lds, hipStream_t s) {
    if (p.max_items <= 0) return hipSuccess;
    if (p.max_items > 0x7FFFFFFF || lds > static_cast<size_t>(kBrickPatchBytes)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KF, dim3(static_cast<unsigned>(p.max_items)), dim3(kBrickThreads), lds, s, p);
    return hipGetLastError();
}

// Gradient tables (cells) have gradient kernels only.
template <int D, int K, bool LASTC, typename T>
hipError_t launch_variant(bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    constexpr bool kSingle = !is_cells<T>();
    if (!grad && !kSingle) return hipErrorInvalidValue;
    if (uniform && p.items) {
        // brick items of the one-pass grouping (2D and 3D grids)
        if constexpr (D >= 2 && D <= kGroupDim) {
            if (!p.item_count || p.cs_dim != 1 || p.cs_pt != static_cast<int64_t>(sizeof(BinRecord) / sizeof(double)))
                return hipErrorInvalidValue;
            const size_t lds = static_cast<size_t>(p.patch_volume) * sizeof(T);
            if (grad) return launch_items<eval_brick_kernel<D, K, true, LASTC, T>>(p, lds, s);
            if constexpr (kSingle) return launch_items<eval_brick_kernel<D, K, false, LASTC, T>>(p, lds, s);
        }
        return hipErrorInvalidValue;
    }
    if (uniform && p.patch_volume > 0) {
        // the patch kernels read the grouping passes' records
        if constexpr (D >= 2 && D <= kGroupDim) {
            if (p.patch_volume > patch_max<T>() || !p.perm || !p.key || p.cs_dim != 1 ||
                p.cs_pt != static_cast<int64_t>(sizeof(BinRecord) / sizeof(double)))
                return hipErrorInvalidValue;
            if (grad) return launch_chunks<eval_patch_kernel<D, K, true, LASTC, T>>(p, s);
            if constexpr (kSingle) return launch_chunks<eval_patch_kernel<D, K, false, LASTC, T>>(p, s);
        }
        return hipErrorInvalidValue;
    }
    if (uniform) {
        if (grad) return launch_chunks<eval_uniform_kernel<D, K, true, LASTC, T>>(p, s);
        if constexpr (kSingle) return launch_chunks<eval_uniform_kernel<D, K, false, LASTC, T>>(p, s);
        return hipErrorInvalidValue;
    }
    if (grad) return launch_chunks<eval_general_kernel<D, K, true, LASTC, T>>(p, s);
    if constexpr (kSingle) return launch_chunks<eval_general_kernel<D, K, false, LASTC, T>>(p, s);
    return hipErrorInvalidValue;
}

template <int D, int K, typename T>
hipError_t launch_one(bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    // contiguous stencil rows need a non-periodic last dimension (stride 1)
    if (p.dim[D - 1].periodic) return launch_variant<D, K, false, T>(grad, uniform, p, s);
    return launch_variant<D, K, true, T>(grad, uniform, p, s);
}

}  // namespace

// Kernels of dimension D with table type T, all orders (one translation
// unit per (T, D) instantiates this).
template <int D, typename T>
hipError_t eval_dispatch(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    switch (order) {
        case 0: return launch_one<D, 0, T>(grad, uniform, p, s);
        case 1: return launch_one<D, 1, T>(grad, uniform, p, s);
        case 2: return launch_one<D, 2, T>(grad, uniform, p, s);
        case 3: return launch_one<D, 3, T>(grad, uniform, p, s);
        case 4: return launch_one<D, 4, T>(grad, uniform, p, s);
        case 5: return launch_one<D, 5, T>(grad, uniform, p, s);
        case 6: return launch_one<D, 6, T>(grad, uniform, p, s);
        case 7: return launch_one<D, 7, T>(grad, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace intp_amd
