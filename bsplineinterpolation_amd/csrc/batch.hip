// batch.hip -- the batched form (intp_batch_*): thousands of independent
// FP64 splines of mixed dimension, order, size and periodicity, built and
// evaluated by a fixed set of persistent launches, optionally captured
// into one hipGraph and replayed, with tile shapes picked by an autotuner
// that times candidates on the batch itself.
//
// Per spline the arithmetic is the single-spline path's (knots and LU on
// the host, InterpolationTemplate.hpp:60-79; line solves last axis first,
// :448-580; evaluation Interpolation.hpp:132-169), so every control point
// equals the one intp_build computes for the same grid and values.
//
// Launch structure of one build + eval:
//   solve pass p (p = 0..2): the axis D-1-p of every spline with D > p,
//     one persistent launch (batch_solve_kernel) over work items of any
//     band width, longest lines first: LDS tiles for contiguous axes (and
//     axes with fewer than 64 adjacent lines), runs of 256 adjacent lines
//     substituted in place for strided axes; the first pass reads the
//     values with the periodic shift;
//   eval: one persistent launch per (dimension, order) class over runs of
//     points of one spline (batch_eval_kernel, batch_eval_d*.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/intp_amd.h"
#include "batch_impl.hpp"
#include "grid.hpp"
#include "runtime.hpp"
#include "solve_impl.hpp"

namespace intp_amd {

namespace {

// Rows per substitution batch in the merged kernels (bounds registers:
// two workgroups per CU).
constexpr int kMergedRows = 4;
// Small tiles (at most kSmallTileBytes of LDS) are solved by one-wave
// workgroups, so that every resident wave substitutes a tile of its own
// (a 256-thread workgroup has one substituting wave and three waiting at
// the barrier); larger tiles and the in-place runs keep 256 threads, whose
// four waves share the loads and stores of a large tile.
constexpr int kSmallSolveThreads = 64;
constexpr int64_t kSmallTileBytes = 20 * 1024;
// Medium tiles (at most kMediumTileBytes: with the 1 KB of line bases,
// four per CU) take two-wave workgroups: at two waves per SIMD four of
// them are resident, so four waves per CU substitute instead of the two
// of two 256-thread workgroups.  A launch allocates its largest item's
// tile for every workgroup, hence a launch per size class.
constexpr int kMediumSolveThreads = 128;
constexpr int64_t kMediumTileBytes = 39 * 1024;
constexpr int kLineThreads = 64;
constexpr int kMaxBatchBand = 6;
// 1D splines of at least this many points are split (INTP_BATCH_SPLIT_LINES)
constexpr int64_t kSplitMinRows = 256;

template <int BW, bool CYC, int NT>
__device__ __forceinline__ void batch_solve_item(const BatchSpline* specs, const BatchSolveItem& w, const double* values, double* ctrl,
                                                 int32_t first_pass, unsigned char* tile_bytes, int64_t* dst_base,
                                                 int64_t* src_base) {
    if ((w.kind >> 4) == kSolveTile) {
        const BatchSpline& b = specs[w.spline];
        const int a = w.axis;
        SolveParams sp{};
        sp.x = ctrl + b.offset;
        sp.n = b.ep.dim[a].n;
        int64_t inner = 1;
        for (int d = a + 1; d < b.dim; ++d) inner *= b.ep.dim[d].n;
        sp.inner = inner;
        sp.nlines = b.total / sp.n;
        sp.lower = b.lower[a];
        sp.upper = b.upper[a];
        sp.diag = b.diag[a];
        sp.bottom = b.bottom[a];
        sp.right = b.right[a];
        sp.lines_per_block = b.lines_per_block[a];
        if (first_pass) {
            sp.src = values + b.offset;
            sp.outer_dims = a;
#pragma unroll
            for (int d = 0; d < kBatchDim; ++d) {
                if (d >= a) break;
                sp.outer_n[d] = b.ep.dim[d].n;
                sp.outer_shift[d] = b.shift[d];
            }
            sp.shift_last = b.shift[a];
        }
        solve_tile<BW, CYC, double, double, kMergedRows, NT>(sp, w.first_line, tile_bytes, dst_base, src_base);
        return;
    }
    // kSolveLanes: one line per lane, in place
    const BatchSpline& b = specs[w.spline];
    const int a = w.axis;
    const int64_t n = b.ep.dim[a].n;
    const int64_t line = w.first_line + threadIdx.x;
    if (line >= b.total / n) return;
    int64_t inner = 1;
    for (int d = a + 1; d < b.dim; ++d) inner *= b.ep.dim[d].n;
    SolveParams sp{};
    sp.lower = b.lower[a];
    sp.upper = b.upper[a];
    sp.diag = b.diag[a];
    sp.bottom = b.bottom[a];
    sp.right = b.right[a];
    const int64_t o = line / inner, ii = line - o * inner;
    substitute<BW, CYC, double, double, kMergedRows>(ctrl + b.offset + o * n * inner + ii, inner, n, sp);
}

// One-dimensional splines: one spline per 64-thread workgroup, longest
// first (list1d).  The wave copies the values into the control-point array
// with the periodic shift (coalesced), then its first lane substitutes the
// line in place.  Runs on its own stream beside the passes of the
// multi-dimensional splines: a 1D line is one long sequential chain, and
// needs a wave, not a workgroup's LDS.
template <int BW, bool CYC>
__device__ __forceinline__ void batch_line(const BatchSpline& b, const double* values, double* ctrl) {
    const int64_t n = b.ep.dim[0].n, shift = b.shift[0];
    double* x = ctrl + b.offset;
    const double* f = values + b.offset;
    if constexpr (BW == 0) {
        // a diagonal system: each row is one division, every lane takes rows
        const double* __restrict__ diag = b.diag[0];
        for (int64_t j = threadIdx.x; j < n; j += kLineThreads) {
            int64_t js = j - shift;
            if (js < 0) js += n;
            x[j] = f[js] / diag[j];
        }
        return;
    }
    for (int64_t j = threadIdx.x; j < n; j += kLineThreads) {
        int64_t js = j - shift;
        if (js < 0) js += n;
        x[j] = f[js];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    SolveParams sp{};
    sp.lower = b.lower[0];
    sp.upper = b.upper[0];
    sp.diag = b.diag[0];
    sp.bottom = b.bottom[0];
    sp.right = b.right[0];
    substitute<BW, CYC, double, double, kMergedRows>(x, 1, n, sp);
}

// Split solve of one long 1D line (INTP_BATCH_SPLIT_LINES).  The first
// H = n - 2*BW - 8 rows are cut into S <= 63 segments, one per lane.  The
// forward sweep is linear in the BW values entering a segment, so each
// lane first substitutes its segment from a zero state (y0), one lane then
// chains the states through the segments (state out = y0's last BW rows +
// gf x state in, staged in LDS), and each lane adds gf x (its state in) to
// its rows.  The last rows, where the cyclic strips and corners act, run
// sequentially with the exact state, the strip sums of the head taken per
// segment.  The backward sweep mirrors this from the right (gb), with the
// corner values as constants.  Same system, different rounding order.
struct SplitShared {
    double state[kLineThreads][kMaxBatchBand];  // state entering each segment
    double y[kLineThreads][kMaxBatchBand];      // a segment's zero-state boundary rows
    double g[kLineThreads][kMaxBatchBand][kMaxBatchBand];  // and their responses
    double vec[2][kMaxBatchBand];               // corners; backward state from the tail
};

template <int BW, bool CYC>
__device__ __forceinline__ void batch_line_split(const BatchSpline& b, const double* values, double* ctrl,
                                                 SplitShared& sh) {
#pragma clang fp contract(off)
    constexpr int kB = BW > 0 ? BW : 1;
    const int64_t n = b.ep.dim[0].n, shift = b.shift[0];
    double* x = ctrl + b.offset;
    const double* f = values + b.offset;
    for (int64_t j = threadIdx.x; j < n; j += kLineThreads) {
        int64_t js = j - shift;
        if (js < 0) js += n;
        x[j] = f[js];
    }
    __syncthreads();
    const int S = b.nseg;
    const int64_t H = n - 2 * BW - 8;
    const int lane = threadIdx.x;
    const int64_t a = static_cast<int64_t>(lane) * H / S, e = static_cast<int64_t>(lane + 1) * H / S;
    SolveParams sp{};
    sp.lower = b.lower[0];
    sp.upper = b.upper[0];
    sp.diag = b.diag[0];
    sp.bottom = b.bottom[0];
    sp.right = b.right[0];
    // global pointers (generic ones read from the record would be flat loads)
    const GlobalPtr<double> __restrict__ bottom = (GlobalPtr<double>)b.bottom[0];
    const GlobalPtr<double> __restrict__ gf = (GlobalPtr<double>)b.gf;
    const GlobalPtr<double> __restrict__ gb = (GlobalPtr<double>)b.gb;
    const int32_t z = lane_zero();
    // forward, zero state entering each segment (the head has no strip rows)
    if (lane < S) {
        LineState<BW, false, double> st0;
#pragma unroll
        for (int m = 0; m <= BW; ++m) st0.hist[m] = st0.nxt[m] = st0.corner[m] = st0.acc[m] = 0.0;
        forward_range<BW, false, double, double, kMergedRows>(x, 1, n, a, e, sp, st0, z);
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            sh.y[lane][q] = x[e - 1 - q];
#pragma unroll
            for (int m = 0; m < BW; ++m) sh.g[lane][q][m] = gf[(e - 1 - q) * BW + m];
        }
    }
    __syncthreads();
    // chain the states through the segments
    if (lane == 0) {
        double h[kB] = {};
        for (int s = 0; s < S; ++s) {
            double out[kB];
#pragma unroll
            for (int q = 0; q < BW; ++q) {
                sh.state[s][q] = h[q];
                double v = sh.y[s][q];
#pragma unroll
                for (int m = 0; m < BW; ++m) v = v + sh.g[s][q][m] * h[m];
                out[q] = v;
            }
#pragma unroll
            for (int m = 0; m < BW; ++m) h[m] = out[m];
        }
    }
    __syncthreads();
    if (lane < S) {
        if (lane > 0) {
            double hin[kB];
#pragma unroll
            for (int m = 0; m < BW; ++m) hin[m] = sh.state[lane][m];
#pragma unroll 4
            for (int64_t j = a; j < e; ++j) {
                double v = x[j];
#pragma unroll
                for (int m = 0; m < BW; ++m) v = v + gf[j * BW + m] * hin[m];
                x[j] = v;
            }
        }
        // strip sums of the head rows, per segment (cyclic matrices)
        if constexpr (CYC) {
            double part[kB] = {};
#pragma unroll 4
            for (int64_t j = a; j < e; ++j) {
                const double v = x[j];
#pragma unroll
                for (int q = 0; q < BW; ++q) part[q] = part[q] - bottom[q * n + j] * v;
            }
#pragma unroll
            for (int q = 0; q < BW; ++q) sh.y[lane][q] = part[q];
        }
    }
    __syncthreads();
    // the last rows, sequentially with the exact state; then the first rows
    // of the backward sweep, which fix the corner values
    if (lane == 0) {
        LineState<BW, CYC, double> st;
#pragma unroll
        for (int m = 0; m <= BW; ++m) st.hist[m] = st.nxt[m] = st.corner[m] = st.acc[m] = 0.0;
#pragma unroll
        for (int m = 0; m < BW; ++m) st.hist[m] = x[H - 1 - m];
        if constexpr (CYC) {
#pragma unroll
            for (int q = 0; q < BW; ++q) {
                double acc = x[n - BW + q];
                for (int s = 0; s < S; ++s) acc = acc + sh.y[s][q];
                st.acc[q] = acc;
            }
        }
        forward_range<BW, CYC, double, double, kMergedRows>(x, 1, n, H, n, sp, st, z);
        backward_range<BW, CYC, double, double, kMergedRows>(x, 1, n, H, n, sp, st, z);
#pragma unroll
        for (int m = 0; m < BW; ++m) {
            sh.vec[0][m] = st.corner[m];
            sh.vec[1][m] = st.nxt[m];  // x[H + m]
        }
    }
    __syncthreads();
    // backward, zero state entering each segment from the right
    if (lane < S) {
        LineState<BW, CYC, double> st0;
#pragma unroll
        for (int m = 0; m <= BW; ++m) st0.hist[m] = st0.nxt[m] = st0.corner[m] = st0.acc[m] = 0.0;
#pragma unroll
        for (int c = 0; c < BW; ++c) st0.corner[c] = sh.vec[0][c];
        backward_range<BW, CYC, double, double, kMergedRows>(x, 1, n, a, e, sp, st0, z);
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            sh.y[lane][q] = x[a + q];
#pragma unroll
            for (int m = 0; m < BW; ++m) sh.g[lane][q][m] = gb[(a + q) * BW + m];
        }
    }
    __syncthreads();
    if (lane == 0) {
        double h[kB];
#pragma unroll
        for (int m = 0; m < BW; ++m) h[m] = sh.vec[1][m];
        for (int s = S - 1; s >= 0; --s) {
            double out[kB];
#pragma unroll
            for (int q = 0; q < BW; ++q) {
                sh.state[s][q] = h[q];
                double v = sh.y[s][q];
#pragma unroll
                for (int m = 0; m < BW; ++m) v = v + sh.g[s][q][m] * h[m];
                out[q] = v;
            }
#pragma unroll
            for (int m = 0; m < BW; ++m) h[m] = out[m];
        }
    }
    __syncthreads();
    if (lane < S) {
        double nin[kB];
#pragma unroll
        for (int m = 0; m < BW; ++m) nin[m] = sh.state[lane][m];
#pragma unroll 4
        for (int64_t j = a; j < e; ++j) {
            double v = x[j];
#pragma unroll
            for (int m = 0; m < BW; ++m) v = v + gb[j * BW + m] * nin[m];
            x[j] = v;
        }
    }
}

// skip_split: the lines with segment responses are solved by
// batch_line_split_kernel instead.
__global__ __launch_bounds__(kLineThreads, 4) void batch_line_kernel(const BatchSpline* specs, const int32_t* list1d,
                                                                  const double* values, double* ctrl,
                                                                  int32_t skip_split) {
    const BatchSpline& b = specs[list1d[blockIdx.x]];
    if (skip_split && b.nseg > 0) return;
    const int kind = 2 * b.band[0] + (b.ep.dim[0].periodic ? 1 : 0);
#define INTP_LINE_CASE(BW, CYC) \
    case 2 * BW + CYC:          \
        batch_line<BW, CYC>(b, values, ctrl); \
        break;
    switch (kind) {
        INTP_LINE_CASE(0, false)
        INTP_LINE_CASE(0, true)
        INTP_LINE_CASE(1, false)
        INTP_LINE_CASE(1, true)
        INTP_LINE_CASE(2, false)
        INTP_LINE_CASE(2, true)
        INTP_LINE_CASE(3, false)
        INTP_LINE_CASE(3, true)
        INTP_LINE_CASE(4, false)
        INTP_LINE_CASE(5, false)
        INTP_LINE_CASE(6, false)
        default: break;  // band widths are checked on the host
    }
#undef INTP_LINE_CASE
}

// One split line per workgroup (list: the splines with segment responses).
__global__ __launch_bounds__(kLineThreads) void batch_line_split_kernel(const BatchSpline* specs, const int32_t* list,
                                                                        const double* values, double* ctrl) {
    __shared__ SplitShared sh;
    const BatchSpline& b = specs[list[blockIdx.x]];
    const int kind = 2 * b.band[0] + (b.ep.dim[0].periodic ? 1 : 0);
#define INTP_LINE_CASE(BW, CYC) \
    case 2 * BW + CYC:          \
        batch_line_split<BW, CYC>(b, values, ctrl, sh); \
        break;
    switch (kind) {
        INTP_LINE_CASE(1, false)
        INTP_LINE_CASE(1, true)
        INTP_LINE_CASE(2, false)
        INTP_LINE_CASE(2, true)
        INTP_LINE_CASE(3, false)
        INTP_LINE_CASE(3, true)
        INTP_LINE_CASE(4, false)
        INTP_LINE_CASE(5, false)
        INTP_LINE_CASE(6, false)
        default: break;  // band widths are checked on the host
    }
#undef INTP_LINE_CASE
}

// One launch per pass: workgroup g solves items g, g + gridDim.x, ...
// (tiles of lines of one spline along one axis, of any band width; the
// items are ordered longest line first, so the long sequential lines
// start at once and the short tiles fill in around them).
template <int NT>
__global__ __launch_bounds__(NT, NT == kTileThreads ? 1 : 2) void batch_solve_kernel(const BatchSpline* specs,
                                                                   const BatchSolveItem* items, int32_t nitems,
                                                                   const double* values, double* ctrl,
                                                                   int32_t first_pass) {
    extern __shared__ __align__(16) unsigned char tile_bytes[];
    __shared__ int64_t dst_base[kMaxTileLines];
    __shared__ int64_t src_base[kMaxTileLines];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BatchSolveItem w = items[it];
#define INTP_SOLVE_CASE(BW, CYC)                                                                                  \
    case 2 * BW + CYC:                                                                                            \
        batch_solve_item<BW, CYC, NT>(specs, w, values, ctrl, first_pass, tile_bytes, dst_base, src_base);        \
        break;
        // periodic band widths are floor(k/2) <= 3, clamped ones k - 1 <= 6
        switch (w.kind & 15) {
            INTP_SOLVE_CASE(0, false)
            INTP_SOLVE_CASE(0, true)
            INTP_SOLVE_CASE(1, false)
            INTP_SOLVE_CASE(1, true)
            INTP_SOLVE_CASE(2, false)
            INTP_SOLVE_CASE(2, true)
            INTP_SOLVE_CASE(3, false)
            INTP_SOLVE_CASE(3, true)
            INTP_SOLVE_CASE(4, false)
            INTP_SOLVE_CASE(5, false)
            INTP_SOLVE_CASE(6, false)
            default: break;  // kinds are checked on the host
        }
#undef INTP_SOLVE_CASE
        __syncthreads();
    }
}

constexpr int kEvalStreams = 4;

// LDS budget of one solve workgroup (two per CU).
constexpr int64_t kBatchTileBudget = 80 * 1024 - 2 * kMaxTileLines * 8;

int64_t tile_fit(int64_t n, int cap) {
    const int64_t row = (n | 1) * 8;
    int lpb = kMaxTileLines;
    while (lpb > 0 && (lpb > cap || lpb * row > kBatchTileBudget)) lpb /= 2;
    return lpb;
}

struct SolveGroup {
    int pass;
    int64_t offset, count;
    size_t lds;
    int threads;  // kSmallSolveThreads, kMediumSolveThreads or kTileThreads
};

struct EvalGroup {
    int dim, order;
    int64_t offset, count;
};

// Tunables (picked by the autotuner).
struct Tuning {
    int tile_cap = 64;           // lines per solve tile, at most
    int solve_blocks_per_cu = 4;  // persistent 256-thread solve workgroups per CU
    int eval_threads = 256;       // points per eval item (= workgroup size)
    int eval_blocks_per_cu = 8;   // persistent eval workgroups per CU
};

}  // namespace

}  // namespace intp_amd

using namespace intp_amd;

struct intp_batch_s {
    int device = 0;
    int count = 0;
    int num_cus = 256;
    std::vector<DimSetup> setups;  // count * kBatchDim
    std::vector<int> dims, orders;
    std::vector<int64_t> offsets;  // count + 1
    std::vector<unsigned char> image;
    unsigned char* tables = nullptr;
    std::vector<BatchSpline> host_specs;
    BatchSpline* specs = nullptr;
    double* ctrl = nullptr;
    // solve work
    std::vector<BatchSolveItem> solve_host;
    BatchSolveItem* solve_items = nullptr;
    int64_t solve_cap = 0;
    std::vector<int32_t> list1d_host;  // the 1D splines, longest first
    int32_t* list1d = nullptr;
    double* split_tables = nullptr;    // segment responses of the split 1D lines
    std::vector<int32_t> split_host;   // the split lines, longest first
    int32_t* split_list = nullptr;
    bool split_ready = false;
    hipStream_t side = nullptr;        // runs the 1D splines beside the passes
    hipStream_t side2 = nullptr;       // the split 1D lines, beside both
    // a pass's solve groups run side by side: all but its last on these
    hipStream_t solve_side[2] = {};
    hipEvent_t solve_fork = nullptr, solve_join[2] = {};
    hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
    // the (dimension, order) eval classes run side by side on these
    hipStream_t eval_side[kEvalStreams] = {};
    hipEvent_t eval_fork = nullptr, eval_join[kEvalStreams] = {};
    std::vector<SolveGroup> solve_groups;
    // eval work for the last point layout
    std::vector<BatchEvalItem> eval_host;
    BatchEvalItem* eval_items = nullptr;
    int64_t eval_cap = 0;
    std::vector<EvalGroup> eval_groups;
    std::vector<int64_t> eval_layout;  // point offsets the eval items were made for
    int eval_threads_used = 0;
    // graph of one build + eval
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* graph_key[4] = {};
    bool graph_split = false;
    std::vector<int64_t> graph_layout;
    Tuning tune;
    bool tuned = false;
    std::string tuning_report = "defaults (not tuned)";
    UseTracker uses;

    ~intp_batch_s() {
        DeviceScope scope(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        release_after(device, tables, uses);
        release_after(device, specs, uses);
        release_after(device, ctrl, uses);
        release_after(device, solve_items, uses);
        release_after(device, list1d, uses);
        release_after(device, split_tables, uses);
        release_after(device, split_list, uses);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (side2) {
            (void)hipStreamSynchronize(side2);
            (void)hipStreamDestroy(side2);
        }
        for (int i = 0; i < 2; ++i) {
            if (solve_side[i]) {
                (void)hipStreamSynchronize(solve_side[i]);
                (void)hipStreamDestroy(solve_side[i]);
            }
            if (solve_join[i]) (void)hipEventDestroy(solve_join[i]);
        }
        if (solve_fork) (void)hipEventDestroy(solve_fork);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (join2) (void)hipEventDestroy(join2);
        for (int i = 0; i < kEvalStreams; ++i) {
            if (eval_side[i]) {
                (void)hipStreamSynchronize(eval_side[i]);
                (void)hipStreamDestroy(eval_side[i]);
            }
            if (eval_join[i]) (void)hipEventDestroy(eval_join[i]);
        }
        if (eval_fork) (void)hipEventDestroy(eval_fork);
        release_after(device, eval_items, uses);
    }
};

namespace {

void drop_graph(intp_batch b);

// (Re)builds the solve work items for the current tuning; uploads them on
// `stream` (the host copy stays alive in the batch).  A captured graph
// bakes in the work-item buffers and their contents, so it is dropped,
// and the upload waits for every launch still reading the old items.
int make_solve_work(intp_batch b, hipStream_t stream) {
    // per spline tile sizes (tiles serve the axes with fewer than 64
    // adjacent lines)
    for (int s = 0; s < b->count; ++s) {
        BatchSpline& bs = b->host_specs[static_cast<size_t>(s)];
        const int D = b->dims[static_cast<size_t>(s)];
        for (int d = 0; d < D; ++d) {
            const int64_t lpb = tile_fit(b->setups[static_cast<size_t>(s) * kBatchDim + d].n, b->tune.tile_cap);
            int64_t inner = 1;
            for (int e = d + 1; e < D; ++e) inner *= b->setups[static_cast<size_t>(s) * kBatchDim + e].n;
            if (lpb == 0 && inner < 64 && D > 1)
                return fail(INTP_E_UNSUPPORTED, "spline " + std::to_string(s) + ": line too long for a batch");
            bs.lines_per_block[d] = static_cast<int32_t>(lpb);
        }
    }
    b->solve_host.clear();
    b->solve_groups.clear();
    // a workgroup's time is about its longest line times the work per row
    auto row_cost = [](const DimSetup& ds) { return int64_t{4} + ds.lu.p * (ds.lu.cyclic ? 4 : 2); };
    auto kind_of = [](const DimSetup& ds, int mode) {
        return static_cast<int16_t>((mode << 4) | (2 * ds.lu.p + (ds.lu.cyclic ? 1 : 0)));
    };
    // per pass: the large tiles with the in-place runs (256-thread
    // workgroups), the medium tiles (two-wave workgroups) and the small
    // tiles (one-wave workgroups), each longest first
    std::vector<std::pair<int64_t, BatchSolveItem>> large_items, medium_items, small_items;
    for (int pass = 0; pass < kBatchDim; ++pass) {
        SolveGroup gl{pass, 0, 0, 0, kTileThreads};
        SolveGroup gm{pass, 0, 0, 0, kMediumSolveThreads};
        SolveGroup gs{pass, 0, 0, 0, kSmallSolveThreads};
        large_items.clear();
        medium_items.clear();
        small_items.clear();
        for (int s = 0; s < b->count; ++s) {
            const int D = b->dims[static_cast<size_t>(s)];
            if (D <= pass) continue;
            const int a = D - 1 - pass;
            const DimSetup& ds = b->setups[static_cast<size_t>(s) * kBatchDim + a];
            if (ds.lu.p > (ds.lu.cyclic ? 3 : kMaxBatchBand))
                return fail(INTP_E_UNSUPPORTED, "band width without a compiled solver");
            const BatchSpline& bs = b->host_specs[static_cast<size_t>(s)];
            const int64_t nlines = bs.total / ds.n;
            int64_t inner = 1;
            for (int d = a + 1; d < D; ++d) inner *= b->setups[static_cast<size_t>(s) * kBatchDim + d].n;
            const int64_t cost = ds.n * row_cost(ds);
            if (D == 1) {
                continue;  // batch_line_kernel, beside the passes
            } else if (inner >= 64) {
                for (int64_t first = 0; first < nlines; first += kTileThreads)
                    large_items.push_back(
                        {cost, BatchSolveItem{s, static_cast<int16_t>(a), kind_of(ds, kSolveLanes), first}});
            } else {
                const int64_t lpb = bs.lines_per_block[a];
                const int64_t bytes = lpb * (ds.n | 1) * 8;
                const bool small = bytes <= kSmallTileBytes, medium = !small && bytes <= kMediumTileBytes;
                auto& to_items = small ? small_items : medium ? medium_items : large_items;
                for (int64_t first = 0; first < nlines; first += lpb)
                    to_items.push_back({cost, BatchSolveItem{s, static_cast<int16_t>(a), kind_of(ds, kSolveTile), first}});
                SolveGroup& to = small ? gs : medium ? gm : gl;
                to.lds = std::max(to.lds, static_cast<size_t>(bytes));
            }
        }
        for (auto* items : {&large_items, &medium_items, &small_items}) {
            SolveGroup& grp = items == &large_items ? gl : items == &medium_items ? gm : gs;
            std::stable_sort(items->begin(), items->end(),
                             [](const auto& x, const auto& y) { return x.first > y.first; });
            grp.offset = static_cast<int64_t>(b->solve_host.size());
            for (const auto& pi : *items) b->solve_host.push_back(pi.second);
            grp.count = static_cast<int64_t>(b->solve_host.size()) - grp.offset;
            if (grp.count > 0) b->solve_groups.push_back(grp);
        }
    }
    drop_graph(b);
    b->uses.order_before(stream);
    const int64_t need = std::max<int64_t>(1, static_cast<int64_t>(b->solve_host.size()));
    if (need > b->solve_cap) {
        release_after(b->device, b->solve_items, b->uses);
        b->solve_items = nullptr;
        HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->solve_items), need * sizeof(BatchSolveItem), stream),
                "device allocation (batch work)");
        b->solve_cap = need;
    }
    HIP_TRY(hipMemcpyAsync(b->specs, b->host_specs.data(), b->host_specs.size() * sizeof(BatchSpline),
                           hipMemcpyHostToDevice, stream),
            "hipMemcpyAsync");
    if (!b->solve_host.empty())
        HIP_TRY(hipMemcpyAsync(b->solve_items, b->solve_host.data(), b->solve_host.size() * sizeof(BatchSolveItem),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return INTP_OK;
}

// Eval work items for a point layout (point_offsets[count + 1]).
int make_eval_work(intp_batch b, const int64_t* point_offsets, hipStream_t stream) {
    const std::vector<int64_t> layout(point_offsets, point_offsets + b->count + 1);
    if (layout == b->eval_layout && b->eval_threads_used == b->tune.eval_threads) return INTP_OK;
    if (layout[0] < 0) return fail(INTP_E_INVALID, "point offsets must start at 0 or more");
    for (int s = 0; s < b->count; ++s)
        if (layout[static_cast<size_t>(s) + 1] < layout[static_cast<size_t>(s)])
            return fail(INTP_E_INVALID, "point offsets must be non-decreasing");
    b->eval_host.clear();
    b->eval_groups.clear();
    const int64_t per = b->tune.eval_threads;
    for (int D = 1; D <= kBatchDim; ++D)
        for (int K = 0; K <= kMaxOrder; ++K) {
            EvalGroup g{D, K, static_cast<int64_t>(b->eval_host.size()), 0};
            for (int s = 0; s < b->count; ++s) {
                if (b->dims[static_cast<size_t>(s)] != D || b->orders[static_cast<size_t>(s)] != K) continue;
                const int64_t p0 = layout[static_cast<size_t>(s)], p1 = layout[static_cast<size_t>(s) + 1];
                for (int64_t p = p0; p < p1; p += per)
                    b->eval_host.push_back({s, static_cast<int32_t>(std::min(per, p1 - p)), p});
            }
            g.count = static_cast<int64_t>(b->eval_host.size()) - g.offset;
            if (g.count > 0) b->eval_groups.push_back(g);
        }
    // a captured graph reads the old items (and maybe the old buffer)
    drop_graph(b);
    b->uses.order_before(stream);
    const int64_t need = std::max<int64_t>(1, static_cast<int64_t>(b->eval_host.size()));
    if (need > b->eval_cap) {
        release_after(b->device, b->eval_items, b->uses);
        b->eval_items = nullptr;
        HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->eval_items), need * sizeof(BatchEvalItem), stream),
                "device allocation (batch work)");
        b->eval_cap = need;
    }
    if (!b->eval_host.empty())
        HIP_TRY(hipMemcpyAsync(b->eval_items, b->eval_host.data(), b->eval_host.size() * sizeof(BatchEvalItem),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    b->eval_layout = layout;
    b->eval_threads_used = b->tune.eval_threads;
    return INTP_OK;
}

// Segment responses of every long 1D line (INTP_BATCH_SPLIT_LINES), from
// the same factors and recurrences as the device: gf[j][m] is row j's
// forward value when only the m-th value entering its segment is 1 (zero
// right-hand side), gb[j][m] the backward one.  Computed and uploaded on
// first use.
int prepare_split(intp_batch b, hipStream_t stream) {
    if (b->split_ready) return INTP_OK;
    std::vector<double> tab;
    std::vector<std::pair<int32_t, size_t>> at;  // spline, offset of its gf (gb follows)
    for (int32_t s : b->list1d_host) {
        const DimSetup& ds = b->setups[static_cast<size_t>(s) * kBatchDim];
        const LineFactors& lu = ds.lu;
        const int BW = lu.p;
        const int64_t n = ds.n, H = n - 2 * BW - 8;
        if (n < kSplitMinRows || BW < 1 || H < 64) continue;
        const int S = static_cast<int>(std::min<int64_t>(63, std::max<int64_t>(2, H / 32)));
        const size_t off = tab.size();
        tab.resize(off + 2 * static_cast<size_t>(H) * BW, 0.0);
        double* gf = tab.data() + off;
        double* gb = gf + static_cast<size_t>(H) * BW;
        for (int sg = 0; sg < S; ++sg) {
            const int64_t a = static_cast<int64_t>(sg) * H / S, e = static_cast<int64_t>(sg + 1) * H / S;
            for (int m = 0; m < BW; ++m) {
                double hist[8] = {}, nx[8] = {};
                hist[m] = 1.0;
                for (int64_t j = a; j < e; ++j) {
                    double v = 0.0;
                    for (int k = BW; k >= 1; --k) v = v - lu.lower[static_cast<size_t>(j * BW + k - 1)] * hist[k - 1];
                    gf[j * BW + m] = v;
                    for (int k = BW; k >= 2; --k) hist[k - 1] = hist[k - 2];
                    hist[0] = v;
                }
                nx[m] = 1.0;
                for (int64_t j = e - 1; j >= a; --j) {
                    double v = 0.0;
                    for (int k = BW; k >= 1; --k) v = v - lu.upper[static_cast<size_t>(j * BW + k - 1)] * nx[k - 1];
                    v = v / lu.diag[static_cast<size_t>(j)];
                    gb[j * BW + m] = v;
                    for (int k = BW; k >= 2; --k) nx[k - 1] = nx[k - 2];
                    nx[0] = v;
                }
            }
        }
        at.push_back({s, off});
        b->split_host.push_back(s);
        b->host_specs[static_cast<size_t>(s)].nseg = S;
    }
    if (!tab.empty()) {
        HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->split_tables), tab.size() * sizeof(double), stream),
                "device allocation (split lines)");
        HIP_TRY(hipMemcpyAsync(b->split_tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
        for (const auto& [s, off] : at) {
            BatchSpline& bs = b->host_specs[static_cast<size_t>(s)];
            const int64_t H = bs.ep.dim[0].n - 2 * bs.band[0] - 8;
            bs.gf = b->split_tables + off;
            bs.gb = bs.gf + H * bs.band[0];
        }
        HIP_TRY(hipMemcpyAsync(b->specs, b->host_specs.data(), b->host_specs.size() * sizeof(BatchSpline),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
        HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->split_list), b->split_host.size() * sizeof(int32_t), stream),
                "device allocation (split lines)");
        HIP_TRY(hipMemcpyAsync(b->split_list, b->split_host.data(), b->split_host.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    b->split_ready = true;
    return INTP_OK;
}

int launch_build(intp_batch b, const double* values, hipStream_t stream, bool split = false) {
    const bool lines = !b->list1d_host.empty();
    if (lines) {
        // fork: the 1D splines on the side stream, joined before returning
        HIP_TRY(hipEventRecord(b->fork, stream), "hipEventRecord");
        HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0), "hipStreamWaitEvent");
        const bool seg = split && !b->split_host.empty();
        if (seg) {
            HIP_TRY(hipStreamWaitEvent(b->side2, b->fork, 0), "hipStreamWaitEvent");
            hipLaunchKernelGGL(batch_line_split_kernel, dim3(static_cast<unsigned>(b->split_host.size())),
                               dim3(kLineThreads), 0, b->side2, b->specs, b->split_list, values, b->ctrl);
            HIP_TRY(hipGetLastError(), "batch split line launch");
            HIP_TRY(hipEventRecord(b->join2, b->side2), "hipEventRecord");
        }
        hipLaunchKernelGGL(batch_line_kernel, dim3(static_cast<unsigned>(b->list1d_host.size())), dim3(kLineThreads),
                           0, b->side, b->specs, b->list1d, values, b->ctrl, seg ? 1 : 0);
        HIP_TRY(hipGetLastError(), "batch line launch");
        HIP_TRY(hipEventRecord(b->join, b->side), "hipEventRecord");
    }
    if (!b->solve_fork) {
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipStreamCreateWithFlags(&b->solve_side[i], hipStreamNonBlocking), "hipStreamCreate");
            HIP_TRY(hipEventCreateWithFlags(&b->solve_join[i], hipEventDisableTiming), "hipEventCreate");
        }
        HIP_TRY(hipEventCreateWithFlags(&b->solve_fork, hipEventDisableTiming), "hipEventCreate");
    }
    auto launch_group = [&](const SolveGroup& g, hipStream_t s) -> hipError_t {
        const int32_t first_pass = g.pass == 0 ? 1 : 0;
        if (g.threads == kSmallSolveThreads) {
            // one wave each: 8 resident per CU at 2 waves per SIMD
            const int blocks = static_cast<int>(std::min<int64_t>(g.count, int64_t{b->num_cus} * 16));
            hipLaunchKernelGGL(batch_solve_kernel<kSmallSolveThreads>, dim3(static_cast<unsigned>(blocks)),
                               dim3(kSmallSolveThreads), g.lds, s, b->specs, b->solve_items + g.offset,
                               static_cast<int32_t>(g.count), values, b->ctrl, first_pass);
        } else if (g.threads == kMediumSolveThreads) {
            // two waves each: 4 resident per CU (LDS and 2 waves per SIMD)
            const int blocks = static_cast<int>(std::min<int64_t>(g.count, int64_t{b->num_cus} * 8));
            hipLaunchKernelGGL(batch_solve_kernel<kMediumSolveThreads>, dim3(static_cast<unsigned>(blocks)),
                               dim3(kMediumSolveThreads), g.lds, s, b->specs, b->solve_items + g.offset,
                               static_cast<int32_t>(g.count), values, b->ctrl, first_pass);
        } else {
            const int blocks =
                static_cast<int>(std::min<int64_t>(g.count, int64_t{b->num_cus} * b->tune.solve_blocks_per_cu));
            hipLaunchKernelGGL(batch_solve_kernel<kTileThreads>, dim3(static_cast<unsigned>(blocks)),
                               dim3(kTileThreads), g.lds, s, b->specs, b->solve_items + g.offset,
                               static_cast<int32_t>(g.count), values, b->ctrl, first_pass);
        }
        return hipGetLastError();
    };
    // a pass's groups run side by side (a spline's lines along one axis
    // are all in one size class, so the groups touch different splines):
    // all but the last on the side streams; the next pass starts after all
    // of them
    for (size_t i = 0; i < b->solve_groups.size();) {
        size_t end = i + 1;
        while (end < b->solve_groups.size() && b->solve_groups[end].pass == b->solve_groups[i].pass) ++end;
        const size_t nside = end - i - 1;  // at most 2: three size classes per pass
        if (nside > 2) return fail(INTP_E_INVALID, "more solve groups in a pass than side streams");
        if (nside > 0) {
            HIP_TRY(hipEventRecord(b->solve_fork, stream), "hipEventRecord");
            for (size_t k = 0; k < nside; ++k) {
                HIP_TRY(hipStreamWaitEvent(b->solve_side[k], b->solve_fork, 0), "hipStreamWaitEvent");
                HIP_TRY(launch_group(b->solve_groups[i + k], b->solve_side[k]), "batch solve launch");
                HIP_TRY(hipEventRecord(b->solve_join[k], b->solve_side[k]), "hipEventRecord");
            }
        }
        HIP_TRY(launch_group(b->solve_groups[end - 1], stream), "batch solve launch");
        for (size_t k = 0; k < nside; ++k) HIP_TRY(hipStreamWaitEvent(stream, b->solve_join[k], 0), "hipStreamWaitEvent");
        i = end;
    }
    if (lines) HIP_TRY(hipStreamWaitEvent(stream, b->join, 0), "hipStreamWaitEvent");
    if (lines && split && !b->split_host.empty()) HIP_TRY(hipStreamWaitEvent(stream, b->join2, 0), "hipStreamWaitEvent");
    return INTP_OK;
}

int launch_eval_work(intp_batch b, const double* coords, double* out, hipStream_t stream) {
    // each class is a small launch: fork them over the side streams so that
    // they run together, and join before returning
    if (!b->eval_fork) {
        HIP_TRY(hipEventCreateWithFlags(&b->eval_fork, hipEventDisableTiming), "hipEventCreate");
        for (int i = 0; i < kEvalStreams; ++i) {
            HIP_TRY(hipStreamCreateWithFlags(&b->eval_side[i], hipStreamNonBlocking), "hipStreamCreate");
            HIP_TRY(hipEventCreateWithFlags(&b->eval_join[i], hipEventDisableTiming), "hipEventCreate");
        }
    }
    const int nside = static_cast<int>(std::min<size_t>(kEvalStreams, b->eval_groups.size()));
    HIP_TRY(hipEventRecord(b->eval_fork, stream), "hipEventRecord");
    for (int i = 0; i < nside; ++i) HIP_TRY(hipStreamWaitEvent(b->eval_side[i], b->eval_fork, 0), "hipStreamWaitEvent");
    int next = 0;
    for (const EvalGroup& g : b->eval_groups) {
        hipStream_t s = b->eval_side[next++ % nside];
        BatchEvalArgs a{b->specs, b->eval_items + g.offset, static_cast<int32_t>(g.count), coords, out};
        const int blocks = static_cast<int>(std::min<int64_t>(g.count, int64_t{b->num_cus} * b->tune.eval_blocks_per_cu));
        hipError_t e = hipErrorNotSupported;
        if (g.dim == 1) e = batch_eval_dispatch_d1(g.order, a, blocks, b->tune.eval_threads, s);
        if (g.dim == 2) e = batch_eval_dispatch_d2(g.order, a, blocks, b->tune.eval_threads, s);
        if (g.dim == 3) e = batch_eval_dispatch_d3(g.order, a, blocks, b->tune.eval_threads, s);
        if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "no batched kernel for this dimension/order");
        HIP_TRY(e, "batch eval launch");
    }
    for (int i = 0; i < nside; ++i) {
        HIP_TRY(hipEventRecord(b->eval_join[i], b->eval_side[i]), "hipEventRecord");
        HIP_TRY(hipStreamWaitEvent(stream, b->eval_join[i], 0), "hipStreamWaitEvent");
    }
    return INTP_OK;
}

void drop_graph(intp_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Times one build + eval with the current tuning (the fastest of 5, ms:
// the candidates differ by a few percent, less than run-to-run noise of a
// median of 3).
float time_run(intp_batch b, const double* values, const double* coords, double* out, hipStream_t stream,
               int* rc, bool split) {
    constexpr int kRuns = 5;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float t[kRuns] = {};
    *rc = launch_build(b, values, stream, split);
    if (*rc == INTP_OK) *rc = launch_eval_work(b, coords, out, stream);  // warm
    for (int r = 0; r < kRuns && *rc == INTP_OK; ++r) {
        (void)hipEventRecord(e0, stream);
        *rc = launch_build(b, values, stream, split);
        if (*rc == INTP_OK) *rc = launch_eval_work(b, coords, out, stream);
        (void)hipEventRecord(e1, stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&t[r], e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return *std::min_element(t, t + kRuns);
}

// Coordinate descent over the tunables, each candidate timed on the
// batch's own build + eval.
int autotune(intp_batch b, const double* values, const double* coords, const int64_t* point_offsets, double* out,
             hipStream_t stream, bool split) {
    std::string report;
    auto try_values = [&](const char* name, int Tuning::*field, std::initializer_list<int> cands) -> int {
        int best = b->tune.*field;
        float best_ms = 1e30f;
        std::string line = std::string(name) + ":";
        for (int v : cands) {
            b->tune.*field = v;
            int rc = make_solve_work(b, stream);
            if (rc == INTP_OK) rc = make_eval_work(b, point_offsets, stream);
            if (rc != INTP_OK) return rc;
            const float ms = time_run(b, values, coords, out, stream, &rc, split);
            if (rc != INTP_OK) return rc;
            line += " " + std::to_string(v) + "=" + std::to_string(ms).substr(0, 6) + "ms";
            if (ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        b->tune.*field = best;
        report += line + " -> " + std::to_string(best) + "; ";
        return INTP_OK;
    };
    // the solve knobs dominate and interact: they are visited twice
    int rc;
    for (int round = 0; round < 2; ++round) {
        if ((rc = try_values("tile_cap", &Tuning::tile_cap, {8, 16, 32, 64})) != INTP_OK) return rc;
        if ((rc = try_values("solve_blocks_per_cu", &Tuning::solve_blocks_per_cu, {1, 2, 4, 8})) != INTP_OK) return rc;
        if (round == 0) {
            if ((rc = try_values("eval_threads", &Tuning::eval_threads, {64, 128, 256})) != INTP_OK) return rc;
            if ((rc = try_values("eval_blocks_per_cu", &Tuning::eval_blocks_per_cu, {2, 4, 8, 16})) != INTP_OK)
                return rc;
        }
    }
    if ((rc = make_solve_work(b, stream)) != INTP_OK) return rc;
    if ((rc = make_eval_work(b, point_offsets, stream)) != INTP_OK) return rc;
    b->tuned = true;
    b->tuning_report = report;
    drop_graph(b);
    return INTP_OK;
}

}  // namespace

extern "C" {

int intp_batch_create(intp_ctx ctx_, int32_t count, const intp_grid_desc* descs, intp_batch* out) {
    if (!ctx_ || !descs || !out || count < 1) return fail(INTP_E_INVALID, "null argument or empty batch");
    const int device = ctx_->device;
    std::unique_ptr<intp_batch_s> b(new intp_batch_s());
    b->device = device;
    b->count = count;
    b->setups.resize(static_cast<size_t>(count) * kBatchDim);
    b->dims.resize(static_cast<size_t>(count));
    b->orders.resize(static_cast<size_t>(count));
    for (int s = 0; s < count; ++s) {
        const intp_grid_desc& g = descs[s];
        if (g.dim < 1 || g.dim > kBatchDim || g.order < 0 || g.order > kMaxOrder)
            return fail(INTP_E_UNSUPPORTED, "spline " + std::to_string(s) + ": no kernels for this dimension/order");
        b->dims[static_cast<size_t>(s)] = g.dim;
        b->orders[static_cast<size_t>(s)] = g.order;
    }
    // host setup (knots, collocation LU) of every dimension, on a thread pool
    std::atomic<int> next{0};
    std::vector<std::string> errors(static_cast<size_t>(count));
    auto worker = [&]() {
        for (int s = next.fetch_add(1); s < count; s = next.fetch_add(1)) {
            const intp_grid_desc& g = descs[s];
            for (int d = 0; d < g.dim; ++d) {
                const std::string err =
                    setup_dimension(g.order, g.periodic[d] != 0, g.n[d], g.xmin[d], g.xmax[d], g.coords[d],
                                    b->setups[static_cast<size_t>(s) * kBatchDim + d]);
                if (!err.empty()) {
                    errors[static_cast<size_t>(s)] = err + " (dimension " + std::to_string(d) + ")";
                    break;
                }
            }
        }
    };
    const int nthreads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    for (int s = 0; s < count; ++s)
        if (!errors[static_cast<size_t>(s)].empty())
            return fail(INTP_E_INVALID, "spline " + std::to_string(s) + ": " + errors[static_cast<size_t>(s)]);

    DeviceScope scope(device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    DeviceRuntime* rt = device_runtime(device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    hipStream_t stream = rt->setup;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        b->num_cus = cus;

    // offsets and the packed table image
    b->offsets.assign(static_cast<size_t>(count) + 1, 0);
    struct Item {
        const void* src;
        size_t bytes;
        size_t at;
    };
    std::vector<Item> items;
    size_t image_bytes = 0;
    auto place = [&](const auto& v) -> size_t {
        const size_t at = image_bytes;
        const size_t bytes = v.size() * sizeof(v[0]);
        items.push_back({v.data(), bytes, at});
        image_bytes += (bytes + 255) / 256 * 256;
        return at;
    };
    struct Where {
        size_t knots, recip, poly, lower, upper, diag, bottom, right;
    };
    std::vector<Where> where(static_cast<size_t>(count) * kBatchDim);
    for (int s = 0; s < count; ++s) {
        int64_t total = 1;
        for (int d = 0; d < b->dims[static_cast<size_t>(s)]; ++d) {
            const DimSetup& ds = b->setups[static_cast<size_t>(s) * kBatchDim + d];
            total *= ds.n;
            Where& w = where[static_cast<size_t>(s) * kBatchDim + d];
            w.knots = place(ds.knots);
            w.recip = place(ds.recip);
            w.poly = place(ds.poly);
            w.lower = place(ds.lu.lower);
            w.upper = place(ds.lu.upper);
            w.diag = place(ds.lu.diag);
            w.bottom = place(ds.lu.bottom);
            w.right = place(ds.lu.right);
        }
        b->offsets[static_cast<size_t>(s) + 1] = b->offsets[static_cast<size_t>(s)] + total;
    }
    const int64_t total = b->offsets.back();
    if (total >= (int64_t{1} << 32)) return fail(INTP_E_UNSUPPORTED, "a batch holds fewer than 2^32 control points");
    b->image.assign(std::max<size_t>(image_bytes, 256), 0);
    for (const Item& it : items)
        if (it.bytes) std::memcpy(b->image.data() + it.at, it.src, it.bytes);
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->tables), b->image.size(), stream), "device allocation (batch)");
    HIP_TRY(hipMemcpyAsync(b->tables, b->image.data(), b->image.size(), hipMemcpyHostToDevice, stream),
            "hipMemcpyAsync");
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->specs), static_cast<size_t>(count) * sizeof(BatchSpline), stream),
            "device allocation (batch)");
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->ctrl), static_cast<size_t>(std::max<int64_t>(total, 1)) * 8,
                           stream),
            "device allocation (batch control points)");
    HIP_TRY(b->uses.record(stream), "hipEventRecord");

    // device descriptors
    b->host_specs.assign(static_cast<size_t>(count), BatchSpline{});
    auto dptr = [&](size_t at, size_t bytes) -> const double* {
        return bytes ? reinterpret_cast<const double*>(b->tables + at) : nullptr;
    };
    for (int s = 0; s < count; ++s) {
        BatchSpline& bs = b->host_specs[static_cast<size_t>(s)];
        const int D = b->dims[static_cast<size_t>(s)], K = b->orders[static_cast<size_t>(s)];
        bs.dim = D;
        bs.order = K;
        bs.offset = b->offsets[static_cast<size_t>(s)];
        bs.total = b->offsets[static_cast<size_t>(s) + 1] - bs.offset;
        EvalParams& ep = bs.ep;
        for (int d = 0; d < D; ++d) {
            const DimSetup& ds = b->setups[static_cast<size_t>(s) * kBatchDim + d];
            const Where& w = where[static_cast<size_t>(s) * kBatchDim + d];
            int64_t stride = 1;
            for (int e = d + 1; e < D; ++e) stride *= b->setups[static_cast<size_t>(s) * kBatchDim + e].n;
            ep.dim[d] = make_dim_params(ds, K, stride, dptr(w.knots, ds.knots.size()), dptr(w.recip, ds.recip.size()),
                                        dptr(w.poly, ds.poly.size()));
            bs.lower[d] = dptr(w.lower, ds.lu.lower.size());
            bs.upper[d] = dptr(w.upper, ds.lu.upper.size());
            bs.diag[d] = dptr(w.diag, ds.lu.diag.size());
            bs.bottom[d] = dptr(w.bottom, ds.lu.bottom.size());
            bs.right[d] = dptr(w.right, ds.lu.right.size());
            bs.shift[d] = ds.periodic ? (K / 2) % ds.n : 0;
            bs.band[d] = ds.lu.p;
        }
        ep.ctrl = b->ctrl + bs.offset;
        ep.cs_dim = 1;
        ep.cs_pt = 3;
        ep.os_out = 0;
        ep.os_pt = 1;
        for (int i = 0; i <= kMaxDim; ++i) ep.slot[i] = -1;
        ep.out_index = 0;
    }
    // the 1D splines, longest first (they run beside the passes)
    for (int s = 0; s < count; ++s)
        if (b->dims[static_cast<size_t>(s)] == 1) {
            const LineFactors& lu = b->setups[static_cast<size_t>(s) * kBatchDim].lu;
            if (lu.p > (lu.cyclic ? 3 : kMaxBatchBand))
                return fail(INTP_E_UNSUPPORTED, "band width without a compiled solver");
            b->list1d_host.push_back(s);
        }
    std::stable_sort(b->list1d_host.begin(), b->list1d_host.end(), [&](int32_t x, int32_t y) {
        return b->setups[static_cast<size_t>(x) * kBatchDim].n > b->setups[static_cast<size_t>(y) * kBatchDim].n;
    });
    if (!b->list1d_host.empty()) {
        HIP_TRY(device_alloc(reinterpret_cast<void**>(&b->list1d), b->list1d_host.size() * sizeof(int32_t), stream),
                "device allocation (batch work)");
        HIP_TRY(hipMemcpyAsync(b->list1d, b->list1d_host.data(), b->list1d_host.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync");
        HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "hipStreamCreate");
        HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
        HIP_TRY(hipEventCreateWithFlags(&b->join, hipEventDisableTiming), "hipEventCreate");
        HIP_TRY(hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking), "hipStreamCreate");
        HIP_TRY(hipEventCreateWithFlags(&b->join2, hipEventDisableTiming), "hipEventCreate");
    }
    const int rc = make_solve_work(b.get(), stream);
    if (rc != INTP_OK) return rc;
    *out = b.release();
    return INTP_OK;
}

int intp_batch_destroy(intp_batch b) {
    delete b;
    return INTP_OK;
}

int intp_batch_info(intp_batch b, int64_t* values_total, int64_t* value_offsets) {
    if (!b) return fail(INTP_E_INVALID, "null batch");
    if (values_total) *values_total = b->offsets.back();
    if (value_offsets) std::memcpy(value_offsets, b->offsets.data(), b->offsets.size() * sizeof(int64_t));
    return INTP_OK;
}

int intp_batch_build(intp_batch b, const double* values, void* stream_) {
    if (!b || !values) return fail(INTP_E_INVALID, "null argument");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    DeviceScope scope(b->device);
    const int rc = launch_build(b, values, stream);
    const hipError_t e = b->uses.record(stream);
    if (rc != INTP_OK) return rc;
    HIP_TRY(e, "hipEventRecord");
    return INTP_OK;
}

int intp_batch_eval(intp_batch b, const double* coords, const int64_t* point_offsets, double* out, void* stream_) {
    if (!b || !coords || !point_offsets || !out) return fail(INTP_E_INVALID, "null argument");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    DeviceScope scope(b->device);
    int rc = make_eval_work(b, point_offsets, stream);
    if (rc != INTP_OK) return rc;
    rc = launch_eval_work(b, coords, out, stream);
    const hipError_t e = b->uses.record(stream);
    if (rc != INTP_OK) return rc;
    HIP_TRY(e, "hipEventRecord");
    return INTP_OK;
}

int intp_batch_run(intp_batch b, const double* values, const double* coords, const int64_t* point_offsets,
                   double* out, uint32_t flags, void* stream_) {
    if (!b || !values || !coords || !point_offsets || !out) return fail(INTP_E_INVALID, "null argument");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    DeviceScope scope(b->device);
    int rc = INTP_OK;
    const bool split = (flags & INTP_BATCH_SPLIT_LINES) != 0;
    if (split && (rc = prepare_split(b, stream)) != INTP_OK) return rc;
    if ((flags & INTP_BATCH_AUTOTUNE) && !b->tuned)
        if ((rc = autotune(b, values, coords, point_offsets, out, stream, split)) != INTP_OK) return rc;
    if ((rc = make_eval_work(b, point_offsets, stream)) != INTP_OK) return rc;
    if ((flags & INTP_BATCH_GRAPH) && stream != nullptr) {
        const void* key[4] = {values, coords, out, stream};
        const bool same = b->exec && std::equal(key, key + 4, b->graph_key) && b->graph_layout == b->eval_layout &&
                          b->graph_split == split;
        if (!same) {
            drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            rc = launch_build(b, values, stream, split);
            if (rc == INTP_OK) rc = launch_eval_work(b, coords, out, stream);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(stream, &g);
            if (rc != INTP_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(ec, "hipStreamEndCapture");
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            std::copy(key, key + 4, b->graph_key);
            b->graph_layout = b->eval_layout;
            b->graph_split = split;
        }
        HIP_TRY(hipGraphLaunch(b->exec, stream), "hipGraphLaunch");
    } else {
        rc = launch_build(b, values, stream, split);
        if (rc == INTP_OK) rc = launch_eval_work(b, coords, out, stream);
        if (rc != INTP_OK) return rc;
    }
    HIP_TRY(b->uses.record(stream), "hipEventRecord");
    return INTP_OK;
}

int intp_batch_tuning(intp_batch b, char* buf, int64_t cap) {
    if (!b || !buf || cap < 1) return fail(INTP_E_INVALID, "null argument");
    const std::string s = "tile_cap=" + std::to_string(b->tune.tile_cap) +
                          " solve_blocks_per_cu=" + std::to_string(b->tune.solve_blocks_per_cu) +
                          " eval_threads=" + std::to_string(b->tune.eval_threads) +
                          " eval_blocks_per_cu=" + std::to_string(b->tune.eval_blocks_per_cu) + " | " +
                          b->tuning_report;
    const size_t n = std::min<size_t>(s.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
    return INTP_OK;
}

int intp_batch_download_ctrl(intp_batch b, int32_t spline, double* host, int64_t count) {
    if (!b || !host) return fail(INTP_E_INVALID, "null argument");
    if (spline < 0 || spline >= b->count) return fail(INTP_E_INVALID, "no such spline");
    const int64_t off = b->offsets[static_cast<size_t>(spline)];
    if (count != b->offsets[static_cast<size_t>(spline) + 1] - off)
        return fail(INTP_E_INVALID, "count must equal the spline's control points");
    DeviceScope scope(b->device);
    DeviceRuntime* rt = device_runtime(b->device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    b->uses.order_before(rt->setup);
    HIP_TRY(hipMemcpyAsync(host, b->ctrl + off, static_cast<size_t>(count) * 8, hipMemcpyDeviceToHost, rt->setup),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(rt->setup), "hipStreamSynchronize");
    return INTP_OK;
}

}  // extern "C"
