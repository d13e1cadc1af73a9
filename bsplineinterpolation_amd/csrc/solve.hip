// solve.hip -- control-point solve kernels and launchers (the device code
// is in solve_impl.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "solve_impl.hpp"

namespace intp_amd {

namespace {

// Tiled launches of at most kWarmLines lines (a long 1D spline, say) keep
// one lane per line busy: registers are plentiful, so they substitute in
// batches of INTP_LONG_BATCH_MULT times the rows, and more of each batch's
// load latency hides behind its arithmetic (same operations, same order).
// Measured on 1D lines of 8192 points: 2x takes the periodic order-6 FP64
// solve from 1.91 to 1.69 ms; 4x spills registers and is slower.  The
// lane-per-line HBM kernel keeps the plain batch (no gain measured).
#ifndef INTP_LONG_BATCH_MULT
#define INTP_LONG_BATCH_MULT 2
#endif
template <int BW, bool CYC>
constexpr int long_batch_rows() {
    return batch_rows<BW, CYC>() * INTP_LONG_BATCH_MULT;
}

template <int BW, bool CYC, typename T, typename C, int B>
__global__ __launch_bounds__(kSolveBlock) void solve_kernel(SolveParams sp) {
    const int64_t line = static_cast<int64_t>(blockIdx.x) * kSolveBlock + threadIdx.x;
    warm_factors<BW, CYC, C, kSolveBlock>(sp, sp.n);  // a few lines: all in block 0
    if (line >= sp.nlines) return;
    const int64_t o = line / sp.inner, ii = line - o * sp.inner;
    substitute<BW, CYC, C, T, B>(static_cast<T*>(sp.x) + o * sp.n * sp.inner + ii, sp.inner, sp.n, sp);
}

template <int BW, bool CYC, typename T, typename C, int B>
__global__ __launch_bounds__(kTileThreads) void solve_tile_kernel(SolveParams sp) {
    extern __shared__ __align__(16) unsigned char tile_bytes[];
    __shared__ int64_t dst_base[kMaxTileLines];
    __shared__ int64_t src_base[kMaxTileLines];
    solve_tile<BW, CYC, T, C, B, kTileThreads, true>(sp, static_cast<int64_t>(blockIdx.x) * sp.lines_per_block,
                                                      tile_bytes, dst_base, src_base);
}

template <int BW, bool CYC, typename T, typename C>
hipError_t launch_lines(const SolveParams& p, hipStream_t s) {
    const bool few = p.nlines <= kWarmLines;
    if (p.lines_per_block > 0) {
        const int lpb = p.lines_per_block;
        const int64_t blocks = (p.nlines + lpb - 1) / lpb;
        const size_t lds = static_cast<size_t>(lpb * tile_stride(p.n)) * sizeof(C);
        if (few)
            hipLaunchKernelGGL((solve_tile_kernel<BW, CYC, T, C, long_batch_rows<BW, CYC>()>),
                               dim3(static_cast<unsigned>(blocks)), dim3(kTileThreads), lds, s, p);
        else
            hipLaunchKernelGGL((solve_tile_kernel<BW, CYC, T, C, batch_rows<BW, CYC>()>),
                               dim3(static_cast<unsigned>(blocks)), dim3(kTileThreads), lds, s, p);
        return hipGetLastError();
    }
    const int64_t blocks = (p.nlines + kSolveBlock - 1) / kSolveBlock;
    hipLaunchKernelGGL((solve_kernel<BW, CYC, T, C, batch_rows<BW, CYC>()>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kSolveBlock), 0, s, p);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(kSolveBlock) void shift_copy_kernel(ShiftParams sp) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kSolveBlock + threadIdx.x;
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
    static_cast<T*>(sp.dst)[dst] = static_cast<T>(sp.src[i]);
}

// LDS a tiled workgroup may use: two workgroups per CU on gfx950
// (160 KB per CU), never more than the device allows one workgroup.
constexpr int64_t kTileLdsBudget = 80 * 1024;

template <int BW, typename T, typename C>
hipError_t launch_bw(bool cyclic, const SolveParams& p, hipStream_t s) {
    if (p.nlines == 0) return hipSuccess;
    if (p.lines_per_block > 0) {
        const int lpb = p.lines_per_block;
        if (lpb > kMaxTileLines || kTileThreads % lpb != 0) return hipErrorInvalidValue;
    } else if (p.src || p.accum) {
        return hipErrorInvalidValue;  // the HBM kernel solves in place only
    }
    return cyclic ? launch_lines<BW, true, T, C>(p, s) : launch_lines<BW, false, T, C>(p, s);
}

template <typename T, typename C>
hipError_t launch_typed(int bw, bool cyclic, const SolveParams& p, hipStream_t s) {
    switch (bw) {
        case 0: return launch_bw<0, T, C>(cyclic, p, s);
        case 1: return launch_bw<1, T, C>(cyclic, p, s);
        case 2: return launch_bw<2, T, C>(cyclic, p, s);
        case 3: return launch_bw<3, T, C>(cyclic, p, s);
        case 4: return launch_bw<4, T, C>(cyclic, p, s);
        case 5: return launch_bw<5, T, C>(cyclic, p, s);
        case 6: return launch_bw<6, T, C>(cyclic, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace

hipError_t launch_solve(int bw, bool cyclic, SolveTypes types, const SolveParams& p, hipStream_t s) {
    switch (types) {
        case SolveTypes::kF64: return launch_typed<double, double>(bw, cyclic, p, s);
        case SolveTypes::kF32: return launch_typed<float, float>(bw, cyclic, p, s);
        case SolveTypes::kF32StoreF64Compute: return launch_typed<float, double>(bw, cyclic, p, s);
        default: return hipErrorInvalidValue;
    }
}

int solve_tile_lines(int64_t n, int compute_bytes) {
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 0;
    const int64_t budget = std::min<int64_t>(kTileLdsBudget, max_lds) - 2 * kMaxTileLines * sizeof(int64_t);
    const int64_t row = tile_stride(n) * static_cast<int64_t>(compute_bytes);
    int lpb = kMaxTileLines;
    while (lpb > 0 && lpb * row > budget) lpb /= 2;
    return lpb;
}

hipError_t launch_shift_copy(const ShiftParams& p, bool dst_f32, hipStream_t s) {
    const int64_t blocks = (p.total + kSolveBlock - 1) / kSolveBlock;
    if (blocks == 0) return hipSuccess;
    if (dst_f32)
        hipLaunchKernelGGL(shift_copy_kernel<float>, dim3(static_cast<unsigned>(blocks)), dim3(kSolveBlock), 0, s, p);
    else
        hipLaunchKernelGGL(shift_copy_kernel<double>, dim3(static_cast<unsigned>(blocks)), dim3(kSolveBlock), 0, s, p);
    return hipGetLastError();
}

}  // namespace intp_amd
