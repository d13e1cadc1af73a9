// span.hpp -- per-dimension span search shared by the evaluation kernels
// and the grouping kernels (device code only).  The grouping keys points
// by the first control point of their stencil, computed by the same
// functions the evaluation uses.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace intp_amd {

// Periodic dimensions: x moved into [lo, lo + period) as the reference
// does (BSpline.hpp:130-136).  Clamped dimensions: with `bad`, flags a
// coordinate outside [lo, hi] (the at() check, Interpolation.hpp:482-491).
__device__ __forceinline__ void wrap_or_check(const DimParams& dp, double& x, int* bad) {
    if (dp.periodic) {
        double r = x - dp.lo;
        if (!(r >= 0.0 && r < dp.period)) r = fmod(r, dp.period);
        x = dp.lo + r + (x < dp.lo ? dp.period : 0.0);
    } else if (bad != nullptr && (x < dp.lo || x > dp.hi)) {
        *bad = 1;
    }
}

// Knot i of a uniform dimension of order k without branches (selects
// only), with the host's operation order (no FMA contraction), so that
// the knots equal the reference's (Interpolation.hpp:322-362).
__device__ __forceinline__ double uknot_sel(const DimParams& dp, int k, int i) {
#pragma clang fp contract(off)
    const double f = dp.xmin + (static_cast<double>(i) - dp.half) * dp.dx;
    if (dp.periodic) return f;
    return i <= k ? dp.xmin : (i >= dp.nt - k - 1 ? dp.xmax : f);
}

// Span of x on a uniform dimension of order k: the reference's hint
// (Interpolation.hpp:277-295), one select-based correction, then a check
// against the exact knots; a lane that still disagrees (NaN, coordinates
// beyond int range) takes the binary search (BSpline.hpp:146-156).
__device__ __forceinline__ int uniform_span(const DimParams& dp, int k, double x) {
    const int last = dp.nt - k - 2;
    double v = (x - dp.lo) * dp.inv_dx - dp.shift;
    v = v > 0.0 ? v : 0.0;
    v = v < static_cast<double>(last) ? v : static_cast<double>(last);
    int mu = min(last, static_cast<int>(ceil(v)) + k);
    const double tl = uknot_sel(dp, k, mu), th = uknot_sel(dp, k, mu + 1);
    mu += (x >= th && mu < last) ? 1 : ((x < tl && mu > k) ? -1 : 0);
    const bool ok = (mu == k || uknot_sel(dp, k, mu) <= x) && (mu == last || x < uknot_sel(dp, k, mu + 1));
    if (!ok) {
        int lo = k + 1, hi = last + 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (x < uknot_sel(dp, k, mid))
                hi = mid;
            else
                lo = mid + 1;
        }
        mu = lo - 1;
    }
    return mu;
}

// Brick of a point: per dimension, the first control point of its
// stencil shifted down to the brick, combined in C order.  The span comes
// from the reference's O(1) hint (Interpolation.hpp:277-295) without the
// exact-knot correction of the evaluation: a point keyed one span off at a
// knot only leaves its brick's patch (the evaluation then reads it from
// HBM); results never depend on the key.
// The counting and scatter passes compute every key twice and rely on
// both giving the same value: no FMA contraction, whatever the caller's
// compilation.
__device__ __forceinline__ uint32_t brick_key_of(const DimParams* dims, int ndim, int order, const BrickGrid& bricks,
                                                 const double* x) {
#pragma clang fp contract(off)
    uint32_t key = 0;
#pragma unroll
    for (int d = 0; d < kGroupDim; ++d) {
        if (d >= ndim) break;
        const DimParams& dp = dims[d];
        double xd = x[d];
        wrap_or_check(dp, xd, nullptr);
        const int last = dp.nt - order - 2;
        double v = (xd - dp.lo) * dp.inv_dx - dp.shift;
        v = v > 0.0 ? v : 0.0;  // NaN -> 0
        v = v < static_cast<double>(last) ? v : static_cast<double>(last);
        int start = min(last, static_cast<int>(ceil(v)) + order) - order;
        if (start >= dp.n) start -= static_cast<int>(dp.n);  // periodic, even order: the last span wraps
        key = key * static_cast<uint32_t>(bricks.nbrick[d]) + (static_cast<uint32_t>(start) >> bricks.brick_shift[d]);
    }
    return key;
}

}  // namespace intp_amd
