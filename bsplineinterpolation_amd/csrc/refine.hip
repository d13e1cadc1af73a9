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
// 2^32 elements).
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void band_matvec_kernel(MatvecParams p) {
    const uint64_t e64 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e64 >= static_cast<uint64_t>(p.total)) return;
    const uint32_t e = static_cast<uint32_t>(e64);
    const uint32_t n = static_cast<uint32_t>(p.n), inner = static_cast<uint32_t>(p.inner);
    const uint32_t t = e / inner, ii = e - t * inner;
    const uint32_t j = t % n;
    const uint32_t base = e - j * inner;
    const TI* __restrict__ x = static_cast<const TI*>(p.x);
    const int w = p.w;
    const double* __restrict__ row = p.colloc + static_cast<uint32_t>(j) * static_cast<uint32_t>(2 * w + 1);
    double acc = 0.0;
    for (int m = -w; m <= w; ++m) {
        int64_t jj = static_cast<int64_t>(j) + m;
        if (p.periodic) {
            if (jj < 0)
                jj += n;
            else if (jj >= n)
                jj -= n;
        } else if (jj < 0 || jj >= static_cast<int64_t>(n)) {
            continue;
        }
        acc += row[m + w] * static_cast<double>(x[base + static_cast<uint32_t>(jj) * inner]);
    }
    if (p.src) {
        // residual: the shifted value minus the product; element
        // (i_0..i_{D-1}) reads source element ((i_d - shift_d) mod n_d)
        uint32_t rem = e;
        uint64_t src = 0, mul = 1;
        for (int d = p.dim - 1; d >= 0; --d) {
            const uint32_t nd = static_cast<uint32_t>(p.dn[d]);
            const uint32_t id = rem % nd;
            rem /= nd;
            int64_t s = static_cast<int64_t>(id) - p.dshift[d];
            if (s < 0) s += nd;
            src += static_cast<uint64_t>(s) * mul;
            mul *= nd;
        }
        acc = p.src[src] - acc;
    }
    static_cast<TO*>(p.y)[e] = static_cast<TO>(acc);
}

__global__ __launch_bounds__(kBlock) void add_f32_kernel(float* __restrict__ c, const float* __restrict__ d,
                                                         int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e < total) c[e] = c[e] + d[e];
}

}  // namespace

hipError_t launch_band_matvec(const MatvecParams& p, bool in_f32, bool out_f32, hipStream_t s) {
    if (p.total >= (int64_t{1} << 32)) return hipErrorInvalidValue;
    const int64_t blocks = (p.total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(blocks)), block(kBlock);
    if (in_f32 && out_f32)
        hipLaunchKernelGGL((band_matvec_kernel<float, float>), grid, block, 0, s, p);
    else if (in_f32)
        hipLaunchKernelGGL((band_matvec_kernel<float, double>), grid, block, 0, s, p);
    else if (out_f32)
        hipLaunchKernelGGL((band_matvec_kernel<double, float>), grid, block, 0, s, p);
    else
        hipLaunchKernelGGL((band_matvec_kernel<double, double>), grid, block, 0, s, p);
    return hipGetLastError();
}

hipError_t launch_add_f32(float* c, const float* d, int64_t total, hipStream_t s) {
    const int64_t blocks = (total + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(add_f32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, c, d, total);
    return hipGetLastError();
}

}  // namespace intp_amd
