// refine.hip -- FP64 residuals for the refined-FP32 build.
//
// The control points c (FP32) of a refined build solve
//     (M_0 (x) M_1 (x) ... (x) M_{D-1}) c = S f
// where M_d is dimension d's collocation matrix (the unfactored matrix the
// reference factors, InterpolationTemplate.hpp:254-446) and S f the values
// with the periodic index shift (InterpolationTemplate.hpp:452-462).  A
// refinement sweep forms r = S f - (M_0 (x) ... ) c in FP64, one axis at a
// time (band mat-vec along an axis: the FP64 counterpart of the
// reference's BandMatrix product, BandMatrix.hpp:68-77, 148-172), solves
// for the correction in FP32 with the LU factors, and adds it to c.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kBlock = 256;

// One output element per thread.  The table is viewed as (outer, n, inner)
// along the applied axis; indices use 32-bit arithmetic (tables below
// 2^32 elements) and multiply-high divisions by the invariant extents.
// The half-width W is a template argument so that the 2W+1 loads of a
// row are all issued before the first product.
// With xcd_remap, each XCD takes a contiguous range of blocks, so that
// the blocks of neighbouring rows (which share 2W of their 2W+1 input
// rows) meet in one L2; otherwise block b goes to XCD b mod 8.
template <int W, typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void band_matvec_kernel(MatvecParams p) {
    const uint64_t blk = p.xcd_remap ? static_cast<uint64_t>(xcd_chunk_index()) : blockIdx.x;
    const uint64_t e64 = blk * kBlock + threadIdx.x;
    if (e64 >= static_cast<uint64_t>(p.total)) return;
    const uint32_t e = static_cast<uint32_t>(e64);
    const uint32_t n = static_cast<uint32_t>(p.n), inner = static_cast<uint32_t>(p.inner);
    const uint32_t t = fast_div(e, p.div_inner);
    const uint32_t j = t - fast_div(t, p.div_n) * n;
    const uint32_t base = e - j * inner;
    const TI* __restrict__ x = static_cast<const TI*>(p.x);
    const double* __restrict__ row = p.colloc + j * static_cast<uint32_t>(2 * W + 1);
    // out-of-range columns of a clamped dimension read element j and count
    // as zero (their matrix entries are zero too)
    double xv[2 * W + 1];
#pragma unroll
    for (int m = -W; m <= W; ++m) {
        int64_t jj = static_cast<int64_t>(j) + m;
        bool ok = true;
        if (p.periodic) {
            jj = jj < 0 ? jj + n : jj >= n ? jj - n : jj;
        } else {
            ok = jj >= 0 && jj < n;
            jj = ok ? jj : j;
        }
        const double v = static_cast<double>(x[base + static_cast<uint32_t>(jj) * inner]);
        xv[m + W] = ok ? v : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m <= 2 * W; ++m) acc += row[m] * xv[m];
    if (p.src) {
        // residual: the shifted value minus the product; element
        // (i_0..i_{D-1}) reads source element ((i_d - shift_d) mod n_d)
        uint32_t rem = e;
        uint64_t src = 0, mul = 1;
        for (int d = p.dim - 1; d >= 0; --d) {
            const uint32_t nd = static_cast<uint32_t>(p.dn[d]);
            const uint32_t qd = fast_div(rem, p.div_dn[d]);
            const uint32_t id = rem - qd * nd;
            rem = qd;
            int64_t s = static_cast<int64_t>(id) - p.dshift[d];
            if (s < 0) s += nd;
            src += static_cast<uint64_t>(s) * mul;
            mul *= nd;
        }
        acc = p.src[src] - acc;
    }
    static_cast<TO*>(p.y)[e] = static_cast<TO>(acc);
}

// Row-run variant for the axes before the last (inner > 1): a thread
// takes R consecutive rows j0..j0+R-1 of one column (o, i) and slides a
// window of R + 2W loaded values along them, so an input element is
// loaded (R + 2W) / R times instead of 2W + 1.  Consecutive threads take
// consecutive i (coalesced).  Each output sums its products in the order
// of band_matvec_kernel, so the results are the same.
// Requires n > 2W: a periodic index then needs at most one wrap.
template <int W, int R, typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void band_matvec_rows_kernel(MatvecParams p) {
    const uint64_t blk = p.xcd_remap ? static_cast<uint64_t>(xcd_chunk_index()) : blockIdx.x;
    const uint64_t w64 = blk * kBlock + threadIdx.x;
    if (w64 >= static_cast<uint64_t>(p.work)) return;
    const uint32_t w = static_cast<uint32_t>(w64);
    const uint32_t n = static_cast<uint32_t>(p.n), inner = static_cast<uint32_t>(p.inner);
    const uint32_t t = fast_div(w, p.div_inner);
    const uint32_t i = w - t * inner;
    const uint32_t o = fast_div(t, p.div_runs);
    const uint32_t j0 = (t - o * static_cast<uint32_t>(p.runs)) * R;
    const uint32_t rows = min(static_cast<uint32_t>(R), n - j0);
    const uint32_t base = o * n * inner + i;
    const TI* __restrict__ x = static_cast<const TI*>(p.x);
    // window entry q holds element j0 - W + q; entries past the last row's
    // stencil, and out-of-range columns of a clamped dimension (whose
    // matrix entries are zero), read element j0 and count as zero
    double xv[R + 2 * W];
#pragma unroll
    for (int q = 0; q < R + 2 * W; ++q) {
        int64_t jj = static_cast<int64_t>(j0) - W + q;
        bool ok = static_cast<uint32_t>(q) < rows + 2 * W;
        if (p.periodic) {
            jj = jj < 0 ? jj + n : jj >= n ? jj - n : jj;
        } else {
            ok = ok && jj >= 0 && jj < n;
        }
        jj = ok ? jj : j0;
        const double v = static_cast<double>(x[base + static_cast<uint32_t>(jj) * inner]);
        xv[q] = ok ? v : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (static_cast<uint32_t>(r) >= rows) break;
        const uint32_t j = j0 + r;
        const uint32_t e = base + j * inner;
        const double* __restrict__ row = p.colloc + j * static_cast<uint32_t>(2 * W + 1);
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m <= 2 * W; ++m) acc += row[m] * xv[r + m];
        if (p.src) {
            uint32_t rem = e;
            uint64_t src = 0, mul = 1;
            for (int d = p.dim - 1; d >= 0; --d) {
                const uint32_t nd = static_cast<uint32_t>(p.dn[d]);
                const uint32_t qd = fast_div(rem, p.div_dn[d]);
                const uint32_t id = rem - qd * nd;
                rem = qd;
                int64_t s = static_cast<int64_t>(id) - p.dshift[d];
                if (s < 0) s += nd;
                src += static_cast<uint64_t>(s) * mul;
                mul *= nd;
            }
            acc = p.src[src] - acc;
        }
        static_cast<TO*>(p.y)[e] = static_cast<TO>(acc);
    }
}

__global__ __launch_bounds__(kBlock) void add_f32_kernel(float* __restrict__ c, const float* __restrict__ d,
                                                         int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e < total) c[e] = c[e] + d[e];
}

// Gradient-table cells of element e: c = c0 + delta in FP64 (exact: the
// correction is far below c0's last bit), the backward differences of
// those sums along each dimension, each rounded once to FP32.
template <int D, typename C>
__global__ __launch_bounds__(kBlock) void refined_cells_kernel(CellParams p) {
    const uint64_t e64 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e64 >= static_cast<uint64_t>(p.total)) return;
    const uint32_t e = static_cast<uint32_t>(e64);
    const float* __restrict__ c0 = p.c0;
    const float* __restrict__ delta = p.delta;
    auto full = [&](uint32_t i) {
        double v = static_cast<double>(c0[i]);
        if (delta) v += static_cast<double>(delta[i]);
        return v;
    };
    const double c = full(e);
    float diff[3] = {0.f, 0.f, 0.f};
    uint32_t rem = e, stride = 1;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const uint32_t nd = static_cast<uint32_t>(p.n[d]);
        const uint32_t q = fast_div(rem, p.div_n[d]);
        const uint32_t id = rem - q * nd;
        rem = q;
        if (id > 0)
            diff[d] = static_cast<float>(c - full(e - stride));
        else if (p.periodic[d])
            diff[d] = static_cast<float>(c - full(e + (nd - 1) * stride));
        stride *= nd;
    }
    C out;
    out.c = static_cast<float>(c);
    out.e0 = diff[0];
    if constexpr (D >= 2) {
        out.e1 = diff[1];
        out.e2 = diff[2];
    }
    static_cast<C*>(p.cells)[e] = out;
}

// c0 = c0 + delta, summed in FP64 and rounded once (the cells' c).
__global__ __launch_bounds__(kBlock) void add_rounded_kernel(float* __restrict__ c, const float* __restrict__ d,
                                                             int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e < total) c[e] = static_cast<float>(static_cast<double>(c[e]) + static_cast<double>(d[e]));
}

constexpr int kRunRows = 8;

template <int W>
hipError_t launch_matvec_width(const MatvecParams& p, bool in_f32, bool out_f32, int64_t blocks, hipStream_t s) {
    const dim3 block(kBlock);
    if (p.runs > 0) {
        const dim3 grid(static_cast<unsigned>((p.work + kBlock - 1) / kBlock));
        if (in_f32 && out_f32)
            hipLaunchKernelGGL((band_matvec_rows_kernel<W, kRunRows, float, float>), grid, block, 0, s, p);
        else if (in_f32)
            hipLaunchKernelGGL((band_matvec_rows_kernel<W, kRunRows, float, double>), grid, block, 0, s, p);
        else if (out_f32)
            hipLaunchKernelGGL((band_matvec_rows_kernel<W, kRunRows, double, float>), grid, block, 0, s, p);
        else
            hipLaunchKernelGGL((band_matvec_rows_kernel<W, kRunRows, double, double>), grid, block, 0, s, p);
        return hipGetLastError();
    }
    const dim3 grid(static_cast<unsigned>(blocks));
    if (in_f32 && out_f32)
        hipLaunchKernelGGL((band_matvec_kernel<W, float, float>), grid, block, 0, s, p);
    else if (in_f32)
        hipLaunchKernelGGL((band_matvec_kernel<W, float, double>), grid, block, 0, s, p);
    else if (out_f32)
        hipLaunchKernelGGL((band_matvec_kernel<W, double, float>), grid, block, 0, s, p);
    else
        hipLaunchKernelGGL((band_matvec_kernel<W, double, double>), grid, block, 0, s, p);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_band_matvec(const MatvecParams& p_in, bool in_f32, bool out_f32, hipStream_t s) {
    if (p_in.total >= (int64_t{1} << 32)) return hipErrorInvalidValue;
    MatvecParams p = p_in;
    p.div_inner = make_fast_div(static_cast<uint32_t>(p.inner));
    p.div_n = make_fast_div(static_cast<uint32_t>(p.n));
    for (int d = 0; d < p.dim && d < kMaxDim; ++d) p.div_dn[d] = make_fast_div(static_cast<uint32_t>(p.dn[d]));
    // the blocks of rows j and j+1 are inner / kBlock apart: under the
    // default round-robin over 8 XCDs they share an L2 only when that
    // distance is a multiple of 8
    p.xcd_remap = p.inner > 1 && (p.inner % kBlock != 0 || (p.inner / kBlock) % 8 != 0) ? 1 : 0;
    const int64_t blocks = (p.total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    // axes before the last take the row-run kernel
    p.runs = 0;
    p.work = 0;
    if (p.inner > 1 && p.n > 2 * static_cast<int64_t>(p.w)) {
        p.runs = static_cast<int32_t>((p.n + kRunRows - 1) / kRunRows);
        p.div_runs = make_fast_div(static_cast<uint32_t>(p.runs));
        p.work = p.total / p.n * p.runs;
    }
    switch (p.w) {
        case 0: return launch_matvec_width<0>(p, in_f32, out_f32, blocks, s);
        case 1: return launch_matvec_width<1>(p, in_f32, out_f32, blocks, s);
        case 2: return launch_matvec_width<2>(p, in_f32, out_f32, blocks, s);
        case 3: return launch_matvec_width<3>(p, in_f32, out_f32, blocks, s);
        case 4: return launch_matvec_width<4>(p, in_f32, out_f32, blocks, s);
        case 5: return launch_matvec_width<5>(p, in_f32, out_f32, blocks, s);
        case 6: return launch_matvec_width<6>(p, in_f32, out_f32, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_refined_cells(const CellParams& p_in, hipStream_t s) {
    if (p_in.total >= (int64_t{1} << 32) || p_in.dim < 1 || p_in.dim > kCellMaxDim) return hipErrorInvalidValue;
    CellParams p = p_in;
    int64_t total = 1;
    for (int d = 0; d < p.dim; ++d) {
        p.div_n[d] = make_fast_div(static_cast<uint32_t>(p.n[d]));
        total *= p.n[d];
    }
    if (total != p.total) return hipErrorInvalidValue;
    const int64_t blocks = (p.total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(blocks)), block(kBlock);
    switch (p.dim) {
        case 1: hipLaunchKernelGGL((refined_cells_kernel<1, Cell1>), grid, block, 0, s, p); break;
        case 2: hipLaunchKernelGGL((refined_cells_kernel<2, Cell3>), grid, block, 0, s, p); break;
        default: hipLaunchKernelGGL((refined_cells_kernel<3, Cell3>), grid, block, 0, s, p); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !p.delta) return e;
    // the cells read c0 around each element: update it only afterwards
    hipLaunchKernelGGL(add_rounded_kernel, grid, block, 0, s, const_cast<float*>(p.c0), p.delta, p.total);
    return hipGetLastError();
}

hipError_t launch_add_f32(float* c, const float* d, int64_t total, hipStream_t s) {
    const int64_t blocks = (total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(add_f32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, c, d, total);
    return hipGetLastError();
}

}  // namespace intp_amd
