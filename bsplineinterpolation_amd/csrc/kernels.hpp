// kernels.hpp -- launch interface between the C ABI (abi.hip) and the
// gfx950 kernels (eval.hip, solve.hip).  Plain structs passed by value
// as kernel arguments.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace intp_amd {

constexpr int kMaxDim = 3;
constexpr int kMaxOrder = 7;

// One dimension of a spline as the evaluation kernels see it.
struct DimParams {
    const double* knots;  // nt knots
    const double* recip;  // nt * order*(order+1)/2 reciprocal knot gaps
    int64_t n;            // control points along the dimension
    int64_t stride;       // element stride of the dimension in the table
    int32_t nt;
    int32_t periodic;
    int32_t uniform;
    int32_t pad_;
    double lo, hi, period;  // range(d) and its length
    double inv_dx;          // 1/dx for the uniform span hint
    double shift;           // hint shift: 1 (periodic) or (order+1)/2
    // uniform dimensions: knot formula and basis-polynomial classes
    double xmin, xmax, dx, half;
    const double* poly;  // nclass * (order+1)^2 coefficients
    int32_t nleft, nright, nclass, pad2_;
};

struct EvalParams {
    DimParams dim[kMaxDim];
    const double* ctrl;
    const double* coords;
    int64_t cs_dim, cs_pt;
    double* out;
    int64_t os_out, os_pt;
    int64_t npts;
    // gradient kernels: output index written for value (slot 0) and for
    // d/dx_d (slot d+1), -1 = not requested
    int32_t slot[kMaxDim + 1];
    // single-output kernels: the derivative order per dimension and the
    // output index
    int32_t order[kMaxDim];
    int32_t out_index;
    int* bad;  // set to 1 if a clamped coordinate is outside range (or null)
    // grouped evaluation: point pt writes to output position
    // perm[pt * perm_stride] (null: position pt)
    const uint32_t* perm;
    int64_t perm_stride;
};

// A query point carried through the grouping passes (32 bytes).
struct BinRecord {
    double x[kMaxDim];
    uint32_t idx;  // position in the caller's arrays
    uint32_t key;  // brick id
};

struct BinDim {
    double lo, period, inv_period, inv_dx;
    int32_t periodic, ncell, brick_shift, nbrick;
};

struct BinParams {
    BinDim dim[kMaxDim];
    int32_t dim_count;
    int32_t nblocks;
    const double* coords;
    int64_t cs_dim, cs_pt;
    int64_t npts;
    const BinRecord* rec_in;
    BinRecord* rec_out;
    const uint32_t* key_in;  // pass 1: keys of rec_in, dense
    uint32_t* key_out;       // pass 0: keys of rec_out, dense
};

// Line solves along one axis of a C-order table viewed as
// (outer, n, inner): line (o, i) holds elements o*n*inner + j*inner + i.
struct SolveParams {
    double* x;  // solved in place
    // First pass only (last axis, inner == 1): read the right-hand sides
    // from `src` with the periodic shift applied, element (i_0..i_{D-1}) of
    // x taking source element ((i_d - shift_d) mod n_d); nullptr: in place.
    const double* src;
    int64_t n, inner, nlines;
    const double* lower;   // n * bw
    const double* upper;   // n * bw
    const double* diag;    // n
    const double* bottom;  // bw * n (cyclic)
    const double* right;   // bw * n (cyclic)
    int32_t outer_dims;            // dims before the solved (last) axis
    int64_t outer_n[kMaxDim];      // their sizes
    int64_t outer_shift[kMaxDim];  // their shifts
    int64_t shift_last;            // shift along the solved axis
    int32_t lines_per_block;       // LDS-tiled kernel: lines per workgroup (0: one lane per line from HBM)
};

struct ShiftParams {
    const double* src;
    double* dst;
    int64_t total;
    int32_t dim;
    int64_t n[kMaxDim];
    int64_t shift[kMaxDim];
};

// Returns hipSuccess or the launch error; hipErrorNotSupported when no
// kernel is compiled for (dim, order).  `uniform`: every dimension is
// uniform and the table has fewer than 2^32 elements (fast path).
hipError_t launch_eval_grad(int dim, int order, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t launch_eval_single(int dim, int order, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t launch_solve(int bw, bool cyclic, const SolveParams& p, hipStream_t s);
// Lines of length n that one LDS-tiled solve workgroup holds on the
// current device (a power of two up to 64), 0 if a line does not fit.
int solve_tile_lines(int64_t n);

// Grouping of query points by brick (bin.hip): scratch sizes and launch.
int64_t bin_chunk_points();
int64_t bin_hist_entries(int64_t npts);
int64_t bin_scan_tiles(int64_t npts);
hipError_t launch_binning(BinParams p, uint32_t* hist, uint32_t* offs, uint32_t* tile_sums, uint32_t* keys,
                          BinRecord* rec_a, BinRecord* rec_b, hipStream_t s);
hipError_t launch_shift_copy(const ShiftParams& p, hipStream_t s);

}  // namespace intp_amd
