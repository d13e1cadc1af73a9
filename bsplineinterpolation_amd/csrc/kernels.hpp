// kernels.hpp -- launch interface between the C ABI (abi.hip) and the
// gfx950 kernels (eval_impl.hpp, solve.hip, bin.hip, ...).  Plain structs passed by value
// as kernel arguments.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace intp_amd {

constexpr int kMaxDim = 6;    // dimensions of a spline
constexpr int kGroupDim = 3;  // dimensions the grouping by brick handles (2 and 3)
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

// Bricks: along dimension d, the points whose stencil starts at control
// point i (i = span - order) belong to brick i >> brick_shift[d]; the key
// of a point is its brick index in C order.  Brick b's stencils read only
// control points [b*B, b*B + B - 1 + order] (B = 1 << brick_shift), wrapped
// on periodic dimensions.
// LDS bytes of one brick patch in the grouped evaluation kernels.
constexpr int kPatchBytes = 24576;

struct BrickGrid {
    int32_t brick_shift[kGroupDim];
    int32_t nbrick[kGroupDim];
};

// A run of grouped points that share one brick (one-pass grouping,
// group.hip): records start .. start + count - 1, brick `bin`.
struct BrickItem {
    uint32_t start, count, bin, pad_;
};

struct EvalParams {
    DimParams dim[kMaxDim];
    const void* ctrl;  // control points: double, float for a refined build, or its gradient table
    const double* coords;
    int64_t cs_dim, cs_pt;
    double* out;
    float* out32;  // set instead of `out`: results stored as FP32
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
    // grouped evaluation with brick patches in LDS (patch_volume > 0):
    // point pt's brick key is key[pt * perm_stride]; each workgroup
    // stages one brick's (B + order)^D control points
    const uint32_t* key;
    BrickGrid bricks;
    int32_t patch_volume;
    // gradient kernels: the D + 1 outputs of a point are contiguous and
    // 16-byte aligned (os_out == 1, slot[s] == s, os_pt even): written as
    // 16-byte stores
    int32_t out_packed;
    // brick-item evaluation (one-pass grouping): items[i] for i <
    // *item_count (at most max_items launched workgroups), each staging
    // its brick's patch_volume elements in dynamic LDS
    const BrickItem* items;
    const uint32_t* item_count;
    int64_t max_items;
};

// A query point carried through the grouping passes (32 bytes).
struct BinRecord {
    double x[kGroupDim];
    uint32_t idx;  // position in the caller's arrays
    uint32_t key;  // brick id
};

struct BinParams {
    DimParams dim[kGroupDim];  // the grid (span search), uniform dimensions only
    BrickGrid bricks;
    int32_t dim_count;
    int32_t order;
    int32_t nblocks;
    const double* coords;
    int64_t cs_dim, cs_pt;
    int64_t npts;
    const BinRecord* rec_in;
    BinRecord* rec_out;
    // brick keys (16 bits) of the pass-0 records, dense: written by the
    // pass-0 scatter, read by the pass-1 histogram
    const uint16_t* key_in;
    uint16_t* key_out;
};

// Evaluation proxies (proxy.hip): create fills first/weights for npts
// points; apply contracts them with a spline's table into out / out32.
//   first[i * dim + d]                      = span - order (before wrap)
//   weights[(i * dim + d) * (order + 1) + m] = value basis
struct ProxyParams {
    DimParams dim[kMaxDim];
    int64_t npts;
    const double* coords;
    int64_t cs_dim, cs_pt;
    int32_t* first;
    double* weights;
    int* bad;
    const void* ctrl;
    double* out;
    float* out32;
    int64_t os_pt;
};
hipError_t launch_proxy_create(int dim, int order, bool uniform, const ProxyParams& p, hipStream_t s);

// The smallest index of a point with a clamped coordinate outside its
// range (the bounds check's offending point), atomically minimized into
// *first (initialized to all ones by the caller).
struct BoundsParams {
    DimParams dim[kMaxDim];
    int32_t ndim;
    const double* coords;
    int64_t cs_dim, cs_pt;
    int64_t npts;
    uint64_t* first;
};
hipError_t launch_first_outside(const BoundsParams& p, hipStream_t s);
// p[0 .. n-1] = value (a kernel, so that it can be captured in a graph).
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, int64_t n, hipStream_t s);
hipError_t launch_proxy_apply(int dim, int order, bool uniform, bool f32, const ProxyParams& p, hipStream_t s);

// Line solves along one axis of a C-order table viewed as
// (outer, n, inner): line (o, i) holds elements o*n*inner + j*inner + i.
// Element types of a line solve: storage (the table) and compute (the
// factors, the LDS tile and the substitution).
enum class SolveTypes { kF64, kF32, kF32StoreF64Compute };

struct SolveParams {
    void* x;  // the table (storage type), solved in place
    // First pass only (last axis, inner == 1): read the right-hand sides
    // from `src` (FP64) with the periodic shift applied, element
    // (i_0..i_{D-1}) of x taking source element ((i_d - shift_d) mod n_d);
    // nullptr: in place.
    const double* src;
    // If set (LDS-tiled kernel only): the solved lines are added into
    // accum (storage type) instead of being stored to x.
    void* accum;
    int64_t n, inner, nlines;
    // factors in the compute type
    const void* lower;   // n * bw
    const void* upper;   // n * bw
    const void* diag;    // n
    const void* bottom;  // bw * n (cyclic)
    const void* right;   // bw * n (cyclic)
    int32_t outer_dims;            // dims before the solved (last) axis
    int64_t outer_n[kMaxDim];      // their sizes
    int64_t outer_shift[kMaxDim];  // their shifts
    int64_t shift_last;            // shift along the solved axis
    int32_t lines_per_block;       // LDS-tiled kernel: lines per workgroup (0: one lane per line from HBM)
};

struct ShiftParams {
    const double* src;
    void* dst;  // double, or float (launch_shift_copy dst_f32)
    int64_t total;
    int32_t dim;
    int64_t n[kMaxDim];
    int64_t shift[kMaxDim];
};

#ifdef __HIPCC__
// Workgroup b runs on XCD b % 8 (round-robin dispatch).  This maps the
// workgroups of each XCD to one contiguous range of chunk indices, so
// that chunks next to each other in memory are processed at the same
// time by CUs sharing one L2.
__device__ __forceinline__ int64_t xcd_chunk_index() {
    constexpr int64_t kXcd = 8;
    int64_t b = blockIdx.x;
    const int64_t per = static_cast<int64_t>(gridDim.x) / kXcd;
    if (b < per * kXcd) b = (b % kXcd) * per + b / kXcd;
    return b;
}
#endif

// Returns hipSuccess or the launch error; hipErrorNotSupported when no
// kernel is compiled for (dim, order).  `uniform`: every dimension is
// uniform and the table has fewer than 2^32 elements (fast path).
hipError_t launch_eval(int dim, int order, bool grad, bool uniform, bool f32, const EvalParams& p, hipStream_t s);
hipError_t launch_eval_grad(int dim, int order, bool uniform, bool f32, const EvalParams& p, hipStream_t s);
hipError_t launch_eval_single(int dim, int order, bool uniform, bool f32, const EvalParams& p, hipStream_t s);
// per table type and dimension (eval_f64_d*.hip, eval_f32_d*.hip; 4 to 6
// dimensions contract the stencil in a rolled loop)
hipError_t eval_dispatch_f64_d1(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f64_d2(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f64_d3(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d1(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d2(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d3(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f64_d4(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f64_d5(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f64_d6(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d4(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d5(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_dispatch_f32_d6(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t launch_solve(int bw, bool cyclic, SolveTypes types, const SolveParams& p, hipStream_t s);
// Lines of length n that one LDS-tiled solve workgroup holds on the
// current device (a power of two up to 64) for compute elements of
// `compute_bytes`, 0 if a line does not fit.
int solve_tile_lines(int64_t n, int compute_bytes);

// One-pass grouping (group.hip): per-block LDS histograms of the brick
// keys over fixed ranges of points, one exclusive scan of the bin-major
// counts, a scatter of 32-byte records to their bins with LDS-atomic
// cursors, and the list of brick items (runs of at most kBrickItemPoints
// records of one brick) for the evaluation kernel.
constexpr int kGroupThreads = 512;
constexpr int kGroupMaxBins = 16384;      // LDS counters of the hist/scatter blocks
constexpr int kBrickThreads = 512;
constexpr int kBrickPointsPerThread = 4;
constexpr int kBrickItemPoints = kBrickThreads * kBrickPointsPerThread;
constexpr int kBrickPatchBytes = 64 * 1024 - 256;  // dynamic LDS of the brick kernel

struct GroupParams {
    DimParams dim[kGroupDim];
    BrickGrid bricks;
    int32_t dim_count;
    int32_t order;
    int32_t nbins;
    int32_t nblocks;    // blocks of the histogram and scatter passes
    int64_t per_block;  // points per block
    const double* coords;
    int64_t cs_dim, cs_pt;
    int64_t npts;
    uint32_t* hist;       // nbins * nblocks counts, bin-major
    uint32_t* offs;       // their exclusive scan
    uint32_t* tile_sums;  // scan scratch
    BinRecord* rec;       // npts records, grouped by bin
    BrickItem* items;     // at most max_items
    uint32_t* item_count;
    int64_t max_items;
};
int64_t group_blocks(int64_t npts, int nbins);
int64_t group_max_items(int64_t npts, int nbins);
hipError_t launch_grouping(GroupParams p, hipStream_t s);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* tile_sums, int64_t n, hipStream_t s);
int64_t scan_tile_sums(int64_t n);

// Grouping of query points by brick (bin.hip): scratch sizes and launch.
int64_t bin_chunk_points();
int64_t bin_hist_entries(int64_t npts);
int64_t bin_scan_tiles(int64_t npts);
hipError_t launch_binning(BinParams p, uint32_t* hist, uint32_t* offs, uint32_t* tile_sums, uint32_t* keys,
                          BinRecord* rec_a, BinRecord* rec_b, hipStream_t s);
hipError_t launch_shift_copy(const ShiftParams& p, bool dst_f32, hipStream_t s);

// y = M x along one axis of a table viewed as (outer, n, inner), M the
// axis' collocation matrix as rows colloc[j*(2w+1) + (m+w)] = M(j, j+m)
// (columns mod n if periodic); with src set, y = (shifted src) - M x.
// Division of 32-bit unsigned integers by an invariant d >= 1 with a
// multiply-high (Granlund & Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, Fig. 4.1): for s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1, q = (t + ((x - t) >> 1)) >> (s - 1)
// with t = mulhi(x, m); d == 1 passes x through.
struct FastDiv {
    uint32_t d = 1, m = 0, s = 0;
};

inline FastDiv make_fast_div(uint32_t d) {
    FastDiv f;
    f.d = d;
    if (d <= 1) return f;
    uint32_t s = 0;
    while ((uint64_t{1} << s) < d) ++s;
    f.s = s;
    f.m = static_cast<uint32_t>(((uint64_t{1} << 32) * ((uint64_t{1} << s) - d)) / d + 1);
    return f;
}

#ifdef __HIPCC__
__device__ __forceinline__ uint32_t fast_div(uint32_t x, const FastDiv& f) {
    if (f.d == 1) return x;
    const uint32_t t = __umulhi(x, f.m);
    return (t + ((x - t) >> 1)) >> (f.s - 1);
}
#endif

struct MatvecParams {
    const void* x;  // float or double
    void* y;        // float or double
    const double* src;
    const double* colloc;
    int32_t w, periodic;
    int64_t n, inner, total;
    int32_t dim;
    int64_t dn[kMaxDim];
    int64_t dshift[kMaxDim];
    // set by launch_band_matvec
    FastDiv div_inner, div_n, div_dn[kMaxDim];
    int32_t xcd_remap;
    // row-run kernel (axes before the last): runs of rows along the axis
    // per column, and the thread count (columns * runs)
    int32_t runs;
    FastDiv div_runs;
    int64_t work;
};
hipError_t launch_band_matvec(const MatvecParams& p, bool in_f32, bool out_f32, hipStream_t s);
hipError_t launch_add_f32(float* c, const float* d, int64_t total, hipStream_t s);

// Gradient table of a refined build: per control point, its FP32 value c
// and its backward differences e_d = c - c[i_d - 1] along each dimension
// (wrapped on periodic dimensions, 0 at i_d = 0 of a clamped one), formed
// in FP64 from the FP64-corrected control points and then rounded.  A
// first partial is the contraction of the differences with the
// nonnegative order-(k-1) weights (de Boor's derivative identity,
// BSpline.hpp:489-520 differentiates the same way), so FP32 rounding of
// the table costs the partials no more than it costs the value.
struct alignas(8) Cell1 {
    float c, e0;
};
struct alignas(16) Cell3 {
    float c, e0, e1, e2;  // 2D tables leave e2 at 0
};
constexpr int kCellMaxDim = 3;  // refined builds of more dimensions keep no gradient table
inline int cell_bytes(int dim) { return dim == 1 ? static_cast<int>(sizeof(Cell1)) : static_cast<int>(sizeof(Cell3)); }
// gradient-table bytes per control point of a refined build (0: none)
inline int gradient_table_bytes(int dim) { return dim <= kCellMaxDim ? cell_bytes(dim) : 0; }

struct CellParams {
    const float* c0;     // FP32 control points
    const float* delta;  // FP32 correction (c = c0 + delta in FP64), or null
    void* cells;         // Cell1 (1D) or Cell3 (2D, 3D)
    int32_t dim;
    int64_t n[kMaxDim];
    int32_t periodic[kMaxDim];
    int64_t total;
    FastDiv div_n[kMaxDim];  // set by launch_refined_cells
};
// cells from (c0, delta); then c0 += delta rounded once from FP64.
// Tables below 2^32 elements.
hipError_t launch_refined_cells(const CellParams& p, hipStream_t s);
// gradient kernels on a gradient table (eval_cells_d*.hip)
hipError_t eval_cells_d1(int order, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_cells_d2(int order, bool uniform, const EvalParams& p, hipStream_t s);
hipError_t eval_cells_d3(int order, bool uniform, const EvalParams& p, hipStream_t s);

}  // namespace intp_amd
