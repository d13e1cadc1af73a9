// solve.hip -- control-point solve on gfx950: every grid line of one axis
// is an independent band (or cyclic band) system sharing one LU
// factorization (InterpolationTemplate.hpp:448-580).  One lane solves one
// line.
//
// Main kernel (solve_tile_kernel): a workgroup copies the lines of up to
// 64 neighbouring lanes into LDS with coalesced bulk loads (runs of one
// line for the contiguous last axis, runs across lines for the strided
// axes), each lane substitutes its line in LDS, and the workgroup writes
// the tile back the same way.  The first pass reads the caller's values
// with the periodic index shift applied (InterpolationTemplate.hpp:452-462),
// so no separate copy precedes it.  Lines too long for LDS fall back to
// solve_kernel, which substitutes in place in HBM.
//
// The substitution is the reference's (BandLU.hpp:121-143, 216-259)
// reordered by rows: each entry receives the same updates in the same
// order, the last `bw` values ride in registers, and the cyclic strips are
// register accumulators (forward) and register-held corner values
// (backward).  Built without FMA contraction and with a true division by
// the pivot, the result is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kTileThreads = 256;
constexpr int kMaxTileLines = 64;

// Elements a lane moves per batch: their loads (values and factor
// coefficients) are issued together, then substituted in order, then
// stored.
constexpr int kBatch = 8;

template <int BW, bool CYC>
struct LineState {
    double hist[BW + 1];    // forward: hist[m-1] = y[j-m]
    double acc[BW + 1];     // forward: rows n-BW+r of a cyclic matrix, strip-updated
    double nxt[BW + 1];     // backward: nxt[m-1] = x[j+m]
    double corner[BW + 1];  // backward: x[n-BW+c]
};

// Forward substitution of rows j0 .. j0+B-1 (unit lower triangle, plus
// the bottom strip of a cyclic matrix).
template <int BW, bool CYC, int B>
__device__ __forceinline__ void forward_rows(double* __restrict__ x, int64_t step, int64_t n, int64_t j0,
                                             const SolveParams& sp, LineState<BW, CYC>& st) {
#pragma clang fp contract(off)
    double xb[B];
    double lc[B][BW + 1];
    double sc[B][BW + 1];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        xb[u] = x[(j0 + u) * step];
#pragma unroll
        for (int m = 0; m < BW; ++m) lc[u][m] = sp.lower[(j0 + u) * BW + m];
        if constexpr (CYC) {
#pragma unroll
            for (int r = 0; r < BW; ++r) sc[u][r] = sp.bottom[r * n + j0 + u];
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 + u;
        double v = xb[u];
        if constexpr (CYC) {
#pragma unroll
            for (int r = 0; r < BW; ++r)
                if (j == n - BW + r) v = st.acc[r];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - lc[u][m - 1] * st.hist[m - 1];
        xb[u] = v;
#pragma unroll
        for (int m = BW; m >= 2; --m) st.hist[m - 1] = st.hist[m - 2];
        if (BW > 0) st.hist[0] = v;
        if constexpr (CYC) {
#pragma unroll
            for (int r = 0; r < BW; ++r)
                if (j < n - BW + r - BW) st.acc[r] = st.acc[r] - sc[u][r] * v;
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) x[(j0 + u) * step] = xb[u];
}

// Backward substitution of rows j0, j0-1, .., j0-B+1 (upper triangle,
// plus the right strip of a cyclic matrix).
template <int BW, bool CYC, int B>
__device__ __forceinline__ void backward_rows(double* __restrict__ x, int64_t step, int64_t n, int64_t j0,
                                              const SolveParams& sp, LineState<BW, CYC>& st) {
#pragma clang fp contract(off)
    double xb[B];
    double uc[B][BW + 1];
    double rc[B][BW + 1];
    double dg[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 - u;
        xb[u] = x[j * step];
#pragma unroll
        for (int m = 0; m < BW; ++m) uc[u][m] = sp.upper[j * BW + m];
        dg[u] = sp.diag[j];
        if constexpr (CYC) {
#pragma unroll
            for (int c = 0; c < BW; ++c) rc[u][c] = sp.right[c * n + j];
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 - u;
        double v = xb[u];
        if constexpr (CYC) {
#pragma unroll
            for (int c = BW - 1; c >= 0; --c)
                if (j < n - BW + c - BW) v = v - rc[u][c] * st.corner[c];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - uc[u][m - 1] * st.nxt[m - 1];
        v = v / dg[u];
        xb[u] = v;
#pragma unroll
        for (int m = BW; m >= 2; --m) st.nxt[m - 1] = st.nxt[m - 2];
        if (BW > 0) st.nxt[0] = v;
        if constexpr (CYC) {
#pragma unroll
            for (int c = 0; c < BW; ++c)
                if (j == n - BW + c) st.corner[c] = v;
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) x[(j0 - u) * step] = xb[u];
}

// Solves one line x[0], x[step], ..., x[(n-1)*step] in place.
template <int BW, bool CYC>
__device__ __forceinline__ void substitute(double* __restrict__ x, int64_t step, int64_t n, const SolveParams& sp) {
    LineState<BW, CYC> st;
#pragma unroll
    for (int m = 0; m <= BW; ++m) st.hist[m] = st.nxt[m] = st.corner[m] = st.acc[m] = 0.0;
    if constexpr (CYC) {
#pragma unroll
        for (int r = 0; r < BW; ++r) st.acc[r] = x[(n - BW + r) * step];
    }
    int64_t j = 0;
    for (; j + kBatch <= n; j += kBatch) forward_rows<BW, CYC, kBatch>(x, step, n, j, sp, st);
    for (; j < n; ++j) forward_rows<BW, CYC, 1>(x, step, n, j, sp, st);
    j = n - 1;
    for (; j - kBatch + 1 >= 0; j -= kBatch) backward_rows<BW, CYC, kBatch>(x, step, n, j, sp, st);
    for (; j >= 0; --j) backward_rows<BW, CYC, 1>(x, step, n, j, sp, st);
}

// Fallback: one lane per line, substituting in HBM.
template <int BW, bool CYC>
__global__ __launch_bounds__(kBlock) void solve_kernel(SolveParams sp) {
    const int64_t line = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (line >= sp.nlines) return;
    const int64_t o = line / sp.inner, ii = line - o * sp.inner;
    substitute<BW, CYC>(sp.x + o * sp.n * sp.inner + ii, sp.inner, sp.n, sp);
}

// Odd LDS row stride (in doubles), so that lanes reading the same column
// of their rows hit distinct banks.
__host__ __device__ __forceinline__ int64_t tile_stride(int64_t n) { return n | 1; }

// One workgroup: lines [blockIdx.x * lpb, + lpb) staged through LDS.
template <int BW, bool CYC>
__global__ __launch_bounds__(kTileThreads) void solve_tile_kernel(SolveParams sp) {
    extern __shared__ double tile[];
    __shared__ int64_t dst_base[kMaxTileLines];
    __shared__ int64_t src_base[kMaxTileLines];
    const int lpb = sp.lines_per_block;
    const int64_t n = sp.n, inner = sp.inner, stride = tile_stride(n);
    const int64_t first = static_cast<int64_t>(blockIdx.x) * lpb;
    const int nl = static_cast<int>(min(static_cast<int64_t>(lpb), sp.nlines - first));
    const int tid = threadIdx.x;
    if (tid < nl) {
        const int64_t line = first + tid;
        const int64_t o = line / inner, ii = line - o * inner;
        dst_base[tid] = o * n * inner + ii;
        if (sp.src) {
            // inner == 1: o runs over the outer dims in C order
            int64_t rem = o, off = 0, mul = n;
            for (int d = sp.outer_dims - 1; d >= 0; --d) {
                const int64_t nd = sp.outer_n[d];
                const int64_t id = rem % nd;
                rem /= nd;
                int64_t w = id - sp.outer_shift[d];
                if (w < 0) w += nd;
                off += w * mul;
                mul *= nd;
            }
            src_base[tid] = off;
        }
    }
    __syncthreads();

    // ---- load ----
    // Contiguous lines: tile element e is (line e / n, element e % n).
    // Strided lines: (line e % lpb, element e / lpb); as 256 % lpb == 0 a
    // thread keeps one line.
    const uint32_t un = static_cast<uint32_t>(n);
    const int lshift = __builtin_ctz(static_cast<unsigned>(lpb));
    const uint32_t count = static_cast<uint32_t>(inner == 1 ? nl * n : lpb * n);
    const uint32_t my_l = static_cast<uint32_t>(tid) & static_cast<uint32_t>(lpb - 1);
    if (sp.src) {
        for (uint32_t e0 = tid; e0 < count; e0 += kTileThreads * kBatch) {
            int64_t from[kBatch];
            uint32_t at[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = min(e0 + u * kTileThreads, count - 1);
                const uint32_t l = e / un, jj = e - l * un;
                int64_t js = static_cast<int64_t>(jj) - sp.shift_last;
                if (js < 0) js += n;
                from[u] = src_base[l] + js;
                at[u] = l * static_cast<uint32_t>(stride) + jj;
            }
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = sp.src[from[u]];
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (e0 + u * kTileThreads < count) tile[at[u]] = v[u];
        }
    } else if (inner == 1) {
        const double* __restrict__ base = sp.x + first * n;
        for (uint32_t e0 = tid; e0 < count; e0 += kTileThreads * kBatch) {
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = base[min(e0 + u * kTileThreads, count - 1)];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * kTileThreads;
                if (e < count) {
                    const uint32_t l = e / un, jj = e - l * un;
                    tile[l * static_cast<uint32_t>(stride) + jj] = v[u];
                }
            }
        }
    } else if (static_cast<int>(my_l) < nl) {
        const double* __restrict__ line = sp.x + dst_base[my_l];
        double* __restrict__ row = tile + my_l * stride;
        for (uint32_t e0 = tid; e0 < count; e0 += kTileThreads * kBatch) {
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t jj = min(e0 + u * kTileThreads, count - 1) >> lshift;
                v[u] = line[static_cast<int64_t>(jj) * inner];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * kTileThreads;
                if (e < count) row[e >> lshift] = v[u];
            }
        }
    }
    __syncthreads();

    if (tid < nl) substitute<BW, CYC>(tile + tid * stride, 1, n, sp);
    __syncthreads();

    // ---- store ----
    if (inner == 1) {
        double* __restrict__ base = sp.x + first * n;
        for (uint32_t e0 = tid; e0 < count; e0 += kTileThreads * kBatch) {
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = min(e0 + u * kTileThreads, count - 1);
                const uint32_t l = e / un, jj = e - l * un;
                v[u] = tile[l * static_cast<uint32_t>(stride) + jj];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (e0 + u * kTileThreads < count) base[e0 + u * kTileThreads] = v[u];
        }
    } else if (static_cast<int>(my_l) < nl) {
        double* __restrict__ line = sp.x + dst_base[my_l];
        const double* __restrict__ row = tile + my_l * stride;
        for (uint32_t e0 = tid; e0 < count; e0 += kTileThreads * kBatch) {
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = row[min(e0 + u * kTileThreads, count - 1) >> lshift];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * kTileThreads;
                if (e < count) line[static_cast<int64_t>(e >> lshift) * inner] = v[u];
            }
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

// LDS a tiled workgroup may use: two workgroups per CU on gfx950
// (160 KB per CU), never more than the device allows one workgroup.
constexpr int64_t kTileLdsBudget = 80 * 1024;

template <int BW>
hipError_t launch_bw(bool cyclic, const SolveParams& p, hipStream_t s) {
    if (p.nlines == 0) return hipSuccess;
    if (p.lines_per_block > 0) {
        const int lpb = p.lines_per_block;
        if (lpb > kMaxTileLines || kTileThreads % lpb != 0) return hipErrorInvalidValue;
        const int64_t blocks = (p.nlines + lpb - 1) / lpb;
        const size_t lds = static_cast<size_t>(lpb * tile_stride(p.n)) * sizeof(double);
        if (cyclic)
            hipLaunchKernelGGL((solve_tile_kernel<BW, true>), dim3(static_cast<unsigned>(blocks)), dim3(kTileThreads),
                               lds, s, p);
        else
            hipLaunchKernelGGL((solve_tile_kernel<BW, false>), dim3(static_cast<unsigned>(blocks)),
                               dim3(kTileThreads), lds, s, p);
        return hipGetLastError();
    }
    if (p.src) return hipErrorInvalidValue;  // the HBM kernel solves in place only
    const int64_t blocks = (p.nlines + kBlock - 1) / kBlock;
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

int solve_tile_lines(int64_t n) {
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 0;
    const int64_t budget = std::min<int64_t>(kTileLdsBudget, max_lds) - 2 * kMaxTileLines * sizeof(int64_t);
    const int64_t row = tile_stride(n) * static_cast<int64_t>(sizeof(double));
    int lpb = kMaxTileLines;
    while (lpb > 0 && lpb * row > budget) lpb /= 2;
    return lpb;
}

hipError_t launch_shift_copy(const ShiftParams& p, hipStream_t s) {
    const int64_t blocks = (p.total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(shift_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

}  // namespace intp_amd
