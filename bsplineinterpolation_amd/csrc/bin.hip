// bin.hip -- groups query points by brick of the control-point grid before
// evaluation, so that the points one workgroup evaluates share a
// cache-resident patch of the table instead of gathering ~20 scattered
// cache lines each from the whole table.
//
// A brick groups the points whose stencils start in the same B^D block of
// control points (BrickGrid).  Two counting-sort passes over the 16-bit
// brick key (low byte, then high byte).  Each pass: per-block LDS histograms (no global atomics), a
// device-wide exclusive scan over the digit-major histogram, and a scatter
// with LDS-atomic ranks.  Chunks are spread over the XCDs in contiguous
// ranges (xcd_chunk_index), so the runs that neighbouring chunks write for
// one digit meet in the same L2.  Ranks inside a block are not stable, which only
// blurs the grouping within one 2048-point block; the evaluation result
// does not depend on the order.  Records carry the original coordinates
// and index, so the evaluation reads them sequentially and writes each
// result to its original position.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "span.hpp"

namespace intp_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 8;
constexpr int kChunk = kThreads * kPerThread;  // points per block
constexpr int kRadix = 256;
constexpr int kScanTile = 4096;  // entries per scan block

__device__ __forceinline__ uint32_t brick_key(const BinParams& p, const double* x) {
    return brick_key_of(p.dim, p.dim_count, p.order, p.bricks, x);
}

__device__ __forceinline__ void load_point(const BinParams& p, int64_t i, double* x) {
#pragma unroll
    for (int d = 0; d < kGroupDim; ++d) x[d] = d < p.dim_count ? p.coords[d * p.cs_dim + i * p.cs_pt] : 0.0;
}

// A record read once, as two 16-byte non-temporal loads.  Streamed
// through L2 at the default priority, the 2 GB of input records evict the
// partly written lines of the digit runs before their neighbouring runs
// complete them: the L2 then writes each output line in pieces (1.8x the
// record bytes in 64-byte write requests) and fetches input lines twice.
__device__ __forceinline__ BinRecord load_record_stream(const BinRecord* src) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    static_assert(sizeof(BinRecord) == 2 * sizeof(u32x4), "BinRecord is two 16-byte halves");
    const u32x4* s = reinterpret_cast<const u32x4*>(src);
    union {
        u32x4 v[2];
        BinRecord r;
    } u;
    u.v[0] = __builtin_nontemporal_load(s);
    u.v[1] = __builtin_nontemporal_load(s + 1);
    return u.r;
}

// ---- exclusive scan of n uint32 (n <= kScanTile * kScanTile) ------------

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wave_sum[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wave) before += wave_sum[w];
        all += wave_sum[w];
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// pass == 0: keys from caller coordinates, digit = low byte
// pass == 1: keys from records, digit = high byte
template <int PASS>
__global__ __launch_bounds__(kThreads) void bin_hist_kernel(BinParams p, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    const int64_t chunk = xcd_chunk_index();
    const int64_t first = chunk * kChunk;
    // every load of the thread is issued before the first digit is formed
    // (a load per LDS atomic leaves one load in flight per thread)
    double x[kPerThread][kGroupDim];
    uint32_t key[kPerThread];
    // pass 1: the thread's kPerThread keys are consecutive (one 16-byte
    // load; the histogram does not depend on which thread counts a key)
    const int64_t kfirst = first + static_cast<int64_t>(threadIdx.x) * kPerThread;
    if constexpr (PASS == 1) {
        static_assert(kPerThread == 8, "one 16-byte load of 16-bit keys");
        if (kfirst + kPerThread <= p.npts) {
            const uint4 v = *reinterpret_cast<const uint4*>(p.key_in + kfirst);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                key[2 * j] = w[j] & 0xFFFFu;
                key[2 * j + 1] = w[j] >> 16;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) key[j] = kfirst + j < p.npts ? p.key_in[kfirst + j] : 0u;
        }
    }
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const int64_t i = first + j * kThreads + threadIdx.x;
        if constexpr (PASS == 0) {
#pragma unroll
            for (int d = 0; d < kGroupDim; ++d) x[j][d] = 0.0;
            if (i < p.npts) load_point(p, i, x[j]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const int64_t i = PASS == 0 ? first + j * kThreads + threadIdx.x : kfirst + j;
        if (i < p.npts) {
            const uint32_t digit = PASS == 0 ? brick_key(p, x[j]) & 0xFFu : (key[j] >> 8) & 0xFFu;
            atomicAdd(&h[digit], 1u);
        }
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * p.nblocks + chunk] = h[threadIdx.x];
}

// Scatter of one chunk: records are first placed in LDS in digit order
// (local rank from an LDS atomic, local start from a block scan of the
// chunk's digit counts), then written out from LDS so that consecutive
// lanes write consecutive slots of one digit's run.
// The chunk goes out in kStageParts rounds of kChunk / kStageParts sorted
// positions, so the LDS stage is that much smaller (more workgroups per
// CU) while each digit's run still spans the whole chunk.
constexpr int kStageParts = 2;

template <int PASS>
__global__ __launch_bounds__(kThreads) void bin_scatter_kernel(BinParams p, const uint32_t* __restrict__ offs) {
    constexpr uint32_t kPart = kChunk / kStageParts;
    __shared__ uint32_t cnt[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t lstart[kRadix];
    __shared__ BinRecord stage[kPart];
    const int64_t chunk = xcd_chunk_index();
    cnt[threadIdx.x] = 0;
    gbase[threadIdx.x] = offs[static_cast<int64_t>(threadIdx.x) * p.nblocks + chunk];
    __syncthreads();
    const int64_t first = chunk * kChunk;
    BinRecord r[kPerThread];
    uint32_t digit[kPerThread], rank[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const int64_t i = first + j * kThreads + threadIdx.x;
        digit[j] = 0;
        rank[j] = 0;
        if (i < p.npts) {
            if constexpr (PASS == 0) {
                load_point(p, i, r[j].x);
                r[j].idx = static_cast<uint32_t>(i);
                r[j].key = brick_key(p, r[j].x);
            } else {
                r[j] = load_record_stream(p.rec_in + i);
            }
            digit[j] = PASS == 0 ? (r[j].key & 0xFFu) : ((r[j].key >> 8) & 0xFFu);
            rank[j] = atomicAdd(&cnt[digit[j]], 1u);
        }
    }
    __syncthreads();
    uint32_t total;
    lstart[threadIdx.x] = block_exclusive_scan(cnt[threadIdx.x], &total);
    __syncthreads();
    const uint32_t n = static_cast<uint32_t>(min<int64_t>(kChunk, p.npts - first));
    uint32_t pos_of[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) pos_of[j] = lstart[digit[j]] + rank[j];
#pragma unroll
    for (int part = 0; part < kStageParts; ++part) {
        const uint32_t lo = part * kPart;
        if (lo >= n) break;  // uniform across the workgroup
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const int64_t i = first + j * kThreads + threadIdx.x;
            if (i < p.npts && pos_of[j] >= lo && pos_of[j] < lo + kPart) stage[pos_of[j] - lo] = r[j];
        }
        __syncthreads();
        const uint32_t m = min(n - lo, kPart);
        for (uint32_t q = threadIdx.x; q < m; q += kThreads) {
            const BinRecord rec = stage[q];
            const uint32_t pos = lo + q;
            const uint32_t d = PASS == 0 ? (rec.key & 0xFFu) : ((rec.key >> 8) & 0xFFu);
            const uint32_t slot = gbase[d] + (pos - lstart[d]);
            p.rec_out[slot] = rec;
            if constexpr (PASS == 0) p.key_out[slot] = static_cast<uint16_t>(rec.key);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                               uint32_t* __restrict__ tile_sums) {
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanTile / kThreads; ++j) {
        const int64_t i = first + j * kThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block: exclusive scan of the tile sums in place (ntiles <= kScanTile).
__global__ __launch_bounds__(kThreads) void scan_tiles_kernel(uint32_t* __restrict__ sums, int64_t ntiles) {
    constexpr int kPer = kScanTile / kThreads;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int64_t i = static_cast<int64_t>(threadIdx.x) * kPer + j;
        v[j] = i < ntiles ? sums[i] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int64_t i = static_cast<int64_t>(threadIdx.x) * kPer + j;
        if (i < ntiles) sums[i] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                              const uint32_t* __restrict__ tile_offs,
                                                              uint32_t* __restrict__ out) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr int kPer = kScanTile / kThreads;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kPer;
    // a thread's kPer entries are 64-byte aligned: whole runs in range move
    // as 16-byte loads and stores
    const bool whole = first + kPer <= n;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; j += 4) {
        if (whole) {
            const u32x4 q = *reinterpret_cast<const u32x4*>(in + first + j);
            v[j] = q.x, v[j + 1] = q.y, v[j + 2] = q.z, v[j + 3] = q.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[j + u] = first + j + u < n ? in[first + j + u] : 0u;
        }
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) s += v[j];
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total) + tile_offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kPer; j += 4) {
        u32x4 q;
        q.x = run, run += v[j];
        q.y = run, run += v[j + 1];
        q.z = run, run += v[j + 2];
        q.w = run, run += v[j + 3];
        if (whole) {
            *reinterpret_cast<u32x4*>(out + first + j) = q;
        } else {
            if (first + j < n) out[first + j] = q.x;
            if (first + j + 1 < n) out[first + j + 1] = q.y;
            if (first + j + 2 < n) out[first + j + 2] = q.z;
            if (first + j + 3 < n) out[first + j + 3] = q.w;
        }
    }
}

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t* tile_sums, int64_t n, hipStream_t s) {
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles > kScanTile) return hipErrorInvalidValue;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kThreads), 0, s, in, n,
                       tile_sums);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(1), dim3(kThreads), 0, s, tile_sums, ntiles);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kThreads), 0, s, in, n,
                       tile_sums, out);
    return hipGetLastError();
}

}  // namespace

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* tile_sums, int64_t n, hipStream_t s) {
    return exclusive_scan(in, out, tile_sums, n, s);
}

int64_t scan_tile_sums(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

int64_t bin_chunk_points() { return kChunk; }

int64_t bin_hist_entries(int64_t npts) { return kRadix * ((npts + kChunk - 1) / kChunk); }

int64_t bin_scan_tiles(int64_t npts) { return (bin_hist_entries(npts) + kScanTile - 1) / kScanTile; }

hipError_t launch_binning(BinParams p, uint32_t* hist, uint32_t* offs, uint32_t* tile_sums, uint32_t* keys,
                          BinRecord* rec_a, BinRecord* rec_b, hipStream_t s) {
    p.nblocks = static_cast<int32_t>((p.npts + kChunk - 1) / kChunk);
    const int64_t nh = static_cast<int64_t>(kRadix) * p.nblocks;
    const dim3 grid(static_cast<unsigned>(p.nblocks)), block(kThreads);
    // pass 0: caller coordinates -> rec_a (and its keys) by low key byte
    p.rec_out = rec_a;
    p.key_out = reinterpret_cast<uint16_t*>(keys);
    hipLaunchKernelGGL(bin_hist_kernel<0>, grid, block, 0, s, p, hist);
    hipError_t e = exclusive_scan(hist, offs, tile_sums, nh, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bin_scatter_kernel<0>, grid, block, 0, s, p, offs);
    // pass 1: rec_a -> rec_b by high key byte
    p.rec_in = rec_a;
    p.key_in = reinterpret_cast<const uint16_t*>(keys);
    p.rec_out = rec_b;
    hipLaunchKernelGGL(bin_hist_kernel<1>, grid, block, 0, s, p, hist);
    e = exclusive_scan(hist, offs, tile_sums, nh, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bin_scatter_kernel<1>, grid, block, 0, s, p, offs);
    return hipGetLastError();
}

}  // namespace intp_amd
