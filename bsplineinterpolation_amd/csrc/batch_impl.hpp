// batch_impl.hpp -- many independent FP64 splines built and evaluated by
// the same launches (the batched form): device-side descriptors, work items
// and the batched evaluation kernel.  Each spline keeps the single-spline
// arithmetic: its lines are solved by solve_tile (solve_impl.hpp) and its
// points evaluated by uniform_point / general_point (eval_impl.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace intp_amd {

constexpr int kBatchDim = 3;  // the batched form takes 1D, 2D and 3D splines

// One spline of a batch, in device memory.
struct BatchSpline {
    EvalParams ep;  // dims, control points (ctrl), value-only output layout
    const double* lower[kBatchDim];
    const double* upper[kBatchDim];
    const double* diag[kBatchDim];
    const double* bottom[kBatchDim];
    const double* right[kBatchDim];
    int64_t offset;  // first value / control point of the spline in the batch arrays
    int64_t total;   // its control points
    int64_t shift[kBatchDim];
    int32_t dim, order;
    int32_t lines_per_block[kBatchDim];  // LDS tile of the solve along each axis
    int32_t band[kBatchDim];             // band width of each axis's LU
    // Split solve of a long 1D line (INTP_BATCH_SPLIT_LINES): nseg segments
    // of the first n - 2*band - 8 rows, and the responses of each segment's
    // rows to the band values entering it, forward (gf) and backward (gb),
    // [row][band].  nseg == 0: the line is solved sequentially.
    const double* gf;
    const double* gb;
    int32_t nseg;
};

// Work of one solve workgroup.  kind & 15 = 2 * band width + cyclic
// selects the substitution; kind >> 4 the mode:
//   kSolveTile: lines first_line .. of `spline` along `axis`, staged
//     through LDS (contiguous axes, and axes with fewer than 64 adjacent
//     lines);
//   kSolveLanes: 256 adjacent lines first_line .. of `spline` along a
//     strided `axis`, one per lane, substituted in place.
enum : int16_t { kSolveTile = 0, kSolveLanes = 1 };

struct BatchSolveItem {
    int32_t spline;
    int16_t axis;
    int16_t kind;
    int64_t first_line;
};

// A run of points of one spline: points [first, first + count) of the
// batch's point arrays.
struct BatchEvalItem {
    int32_t spline;
    int32_t count;
    int64_t first;
};

struct BatchEvalArgs {
    const BatchSpline* specs;
    const BatchEvalItem* items;
    int32_t nitems;
    const double* coords;  // [points][3] (x, y, z; unused dimensions ignored)
    double* out;           // [points]
};

hipError_t batch_eval_dispatch_d1(int order, const BatchEvalArgs& a, int blocks, int threads, hipStream_t s);
hipError_t batch_eval_dispatch_d2(int order, const BatchEvalArgs& a, int blocks, int threads, hipStream_t s);
hipError_t batch_eval_dispatch_d3(int order, const BatchEvalArgs& a, int blocks, int threads, hipStream_t s);

#ifdef INTP_BATCH_EVAL_KERNELS

namespace {

// Persistent: each workgroup takes items blockIdx.x, + gridDim.x, ...; the
// item's spline parameters are copied into LDS, then one thread per point.
template <int D, int K>
__global__ __launch_bounds__(256) void batch_eval_kernel(BatchEvalArgs a) {
    __shared__ EvalParams sp;
    for (int it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const BatchEvalItem w = a.items[it];
        const BatchSpline* __restrict__ b = a.specs + w.spline;
        {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(&b->ep);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&sp);
            for (int i = threadIdx.x; i < static_cast<int>(sizeof(EvalParams) / 4); i += blockDim.x) dst[i] = src[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            sp.coords = a.coords;
            sp.out = a.out;
        }
        const bool uniform = sp.dim[0].uniform && (D < 2 || sp.dim[1].uniform) && (D < 3 || sp.dim[2].uniform) &&
                             sp.dim[0].poly != nullptr;
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < w.count) {
            const int64_t pt = w.first + threadIdx.x;
            const bool lastc = !sp.dim[D - 1].periodic;
            if (uniform) {
                if (lastc)
                    uniform_point<D, K, false, true, double>(sp, pt);
                else
                    uniform_point<D, K, false, false, double>(sp, pt);
            } else {
                if (lastc)
                    general_point<D, K, false, true, double>(sp, pt);
                else
                    general_point<D, K, false, false, double>(sp, pt);
            }
        }
        __syncthreads();
    }
}

template <int D>
hipError_t batch_eval_dispatch(int order, const BatchEvalArgs& a, int blocks, int threads, hipStream_t s) {
    if (a.nitems == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(blocks)), block(static_cast<unsigned>(threads));
    switch (order) {
        case 0: hipLaunchKernelGGL((batch_eval_kernel<D, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((batch_eval_kernel<D, 1>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((batch_eval_kernel<D, 2>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((batch_eval_kernel<D, 3>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((batch_eval_kernel<D, 4>), grid, block, 0, s, a); break;
        case 5: hipLaunchKernelGGL((batch_eval_kernel<D, 5>), grid, block, 0, s, a); break;
        case 6: hipLaunchKernelGGL((batch_eval_kernel<D, 6>), grid, block, 0, s, a); break;
        case 7: hipLaunchKernelGGL((batch_eval_kernel<D, 7>), grid, block, 0, s, a); break;
        default: return hipErrorNotSupported;
    }
    return hipGetLastError();
}

}  // namespace

#endif  // INTP_BATCH_EVAL_KERNELS

}  // namespace intp_amd
