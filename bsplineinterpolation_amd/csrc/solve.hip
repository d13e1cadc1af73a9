// solve.hip -- control-point solve on gfx950: every grid line of one axis
// is an independent band (or cyclic band) system sharing one LU
// factorization (InterpolationTemplate.hpp:448-580).  One lane per line;
// consecutive lanes take consecutive lines, so for every axis but the last
// a wavefront's loads at step j are one contiguous 512-byte run.
//
// The substitution is the reference's (BandLU.hpp:121-143, 216-259)
// reordered by rows: each entry receives the same updates in the same
// order, the last `bw` values ride in registers, and the cyclic strips are
// register accumulators (forward) and register-held corner values
// (backward).  Built without FMA contraction and with a true division by
// the pivot, the result is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kBlock = 256;

template <int BW, bool CYC>
__global__ __launch_bounds__(kBlock) void solve_kernel(SolveParams sp) {
#pragma clang fp contract(off)
    const int64_t line = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (line >= sp.nlines) return;
    const int64_t n = sp.n, inner = sp.inner;
    const int64_t o = line / inner, ii = line - o * inner;
    double* __restrict__ x = sp.x + o * n * inner + ii;
    const double* __restrict__ L = sp.lower;
    const double* __restrict__ U = sp.upper;
    const double* __restrict__ Dg = sp.diag;

    // ---- forward: unit lower triangle (+ bottom strip) ----------------
    double hist[BW + 1];  // hist[m-1] = y[j-m]
#pragma unroll
    for (int m = 0; m <= BW; ++m) hist[m] = 0.0;
    double acc[BW + 1];  // rows n-BW+r of a cyclic matrix, strip-updated
    if constexpr (CYC) {
#pragma unroll
        for (int r = 0; r < BW; ++r) acc[r] = x[(n - BW + r) * inner];
    }
    for (int64_t j = 0; j < n; ++j) {
        double v = x[j * inner];
        if constexpr (CYC) {
#pragma unroll
            for (int r = 0; r < BW; ++r)
                if (j == n - BW + r) v = acc[r];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - L[j * BW + (m - 1)] * hist[m - 1];
        x[j * inner] = v;
#pragma unroll
        for (int m = BW; m >= 2; --m) hist[m - 1] = hist[m - 2];
        if (BW > 0) hist[0] = v;
        if constexpr (CYC) {
            const double* __restrict__ S = sp.bottom;
#pragma unroll
            for (int r = 0; r < BW; ++r)
                if (j < n - BW + r - BW) acc[r] = acc[r] - S[r * n + j] * v;
        }
    }

    // ---- backward: upper triangle (+ right strip) ----------------------
    double nxt[BW + 1];  // nxt[m-1] = x[j+m]
#pragma unroll
    for (int m = 0; m <= BW; ++m) nxt[m] = 0.0;
    double corner[BW + 1];  // x[n-BW+c]
#pragma unroll
    for (int c = 0; c <= BW; ++c) corner[c] = 0.0;
    for (int64_t j = n - 1; j >= 0; --j) {
        double v = x[j * inner];
        if constexpr (CYC) {
            const double* __restrict__ Rs = sp.right;
#pragma unroll
            for (int c = BW - 1; c >= 0; --c)
                if (j < n - BW + c - BW) v = v - Rs[c * n + j] * corner[c];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - U[j * BW + (m - 1)] * nxt[m - 1];
        v = v / Dg[j];
        x[j * inner] = v;
#pragma unroll
        for (int m = BW; m >= 2; --m) nxt[m - 1] = nxt[m - 2];
        if (BW > 0) nxt[0] = v;
        if constexpr (CYC) {
#pragma unroll
            for (int c = 0; c < BW; ++c)
                if (j == n - BW + c) corner[c] = v;
        }
    }
}

__global__ __launch_bounds__(kBlock) void shift_copy_kernel(ShiftParams sp) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= sp.total) return;
    int64_t rem = i, dst = 0, mul = 1;
    for (int d = sp.dim - 1; d >= 0; --d) {
        const int64_t n = sp.n[d];
        const int64_t id = rem % n;
        rem /= n;
        int64_t w = id + sp.shift[d];
        if (w >= n) w -= n;
        dst += w * mul;
        mul *= n;
    }
    sp.dst[dst] = sp.src[i];
}

template <int BW>
hipError_t launch_bw(bool cyclic, const SolveParams& p, hipStream_t s) {
    const int64_t blocks = (p.nlines + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    if (cyclic)
        hipLaunchKernelGGL((solve_kernel<BW, true>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    else
        hipLaunchKernelGGL((solve_kernel<BW, false>), dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_solve(int bw, bool cyclic, const SolveParams& p, hipStream_t s) {
    switch (bw) {
        case 0: return launch_bw<0>(cyclic, p, s);
        case 1: return launch_bw<1>(cyclic, p, s);
        case 2: return launch_bw<2>(cyclic, p, s);
        case 3: return launch_bw<3>(cyclic, p, s);
        case 4: return launch_bw<4>(cyclic, p, s);
        case 5: return launch_bw<5>(cyclic, p, s);
        case 6: return launch_bw<6>(cyclic, p, s);
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_shift_copy(const ShiftParams& p, hipStream_t s) {
    const int64_t blocks = (p.total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(shift_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

}  // namespace intp_amd
