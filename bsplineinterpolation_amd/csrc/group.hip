// group.hip -- one-pass grouping of query points by brick of the control
// point table, for the brick-item evaluation kernel.
//
// Bricks are large (their patch of (B + order)^D control points fills up
// to 64 KB of LDS), so there are few of them (at most kGroupMaxBins) and
// every block of the histogram and scatter passes keeps one counter per
// brick in LDS.  The passes:
//  1. histogram: block b counts the keys of its fixed range of points;
//     counts go out bin-major (bin * nblocks + b);
//  2. exclusive scan of the bin-major counts: the first slot of every
//     (bin, block) pair, so bins are contiguous in the output;
//  3. scatter: block b re-reads its points, recomputes each key, takes a
//     slot from its LDS cursor of that bin and writes the 32-byte record
//     (coordinates, caller index, key) there;
//  4. items: one block cuts every bin into runs of at most
//     kBrickItemPoints records.
// The points cross HBM twice (coordinates read by passes 1 and 3, records
// written once) instead of the four times of two radix passes.  Order
// inside a bin depends on LDS atomics; evaluation results do not.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "span.hpp"

namespace intp_amd {

namespace {

constexpr int kBatch = 4;  // points in flight per thread

__device__ __forceinline__ void load_coords(const GroupParams& p, int64_t i, double* x) {
#pragma unroll
    for (int d = 0; d < kGroupDim; ++d) x[d] = d < p.dim_count ? p.coords[d * p.cs_dim + i * p.cs_pt] : 0.0;
}

__global__ __launch_bounds__(kGroupThreads) void group_hist_kernel(GroupParams p) {
    extern __shared__ uint32_t h[];
    for (int b = threadIdx.x; b < p.nbins; b += kGroupThreads) h[b] = 0;
    __syncthreads();
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * p.per_block;
    const int64_t hi = min(p.npts, lo + p.per_block);
    for (int64_t base = lo; base < hi; base += kGroupThreads * kBatch) {
        double x[kBatch][kGroupDim];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = base + j * kGroupThreads + threadIdx.x;
            if (i < hi) load_coords(p, i, x[j]);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = base + j * kGroupThreads + threadIdx.x;
            if (i < hi) atomicAdd(&h[brick_key_of(p.dim, p.dim_count, p.order, p.bricks, x[j])], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < p.nbins; b += kGroupThreads)
        p.hist[static_cast<int64_t>(b) * p.nblocks + blockIdx.x] = h[b];
}

__device__ __forceinline__ void store_record(BinRecord* dst, const BinRecord& r) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    union {
        BinRecord r;
        u32x4 v[2];
    } u;
    u.r = r;
    u32x4* d = reinterpret_cast<u32x4*>(dst);
    d[0] = u.v[0];
    d[1] = u.v[1];
}

__global__ __launch_bounds__(kGroupThreads) void group_scatter_kernel(GroupParams p) {
    extern __shared__ uint32_t cursor[];
    for (int b = threadIdx.x; b < p.nbins; b += kGroupThreads)
        cursor[b] = p.offs[static_cast<int64_t>(b) * p.nblocks + blockIdx.x];
    __syncthreads();
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * p.per_block;
    const int64_t hi = min(p.npts, lo + p.per_block);
    for (int64_t base = lo; base < hi; base += kGroupThreads * kBatch) {
        BinRecord r[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = base + j * kGroupThreads + threadIdx.x;
            if (i < hi) load_coords(p, i, r[j].x);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = base + j * kGroupThreads + threadIdx.x;
            if (i < hi) {
                r[j].idx = static_cast<uint32_t>(i);
                r[j].key = brick_key_of(p.dim, p.dim_count, p.order, p.bricks, r[j].x);
                const uint32_t slot = atomicAdd(&cursor[r[j].key], 1u);
                if (slot < p.npts) store_record(p.rec + slot, r[j]);
            }
        }
    }
}

// One block: bin b holds records [start(b), start(b + 1)), start(b) being
// the scanned offset of (b, block 0); each bin becomes ceil(count / M)
// items, numbered in bin order.
constexpr int kItemThreads = 1024;

__global__ __launch_bounds__(kItemThreads) void group_items_kernel(GroupParams p) {
    __shared__ uint32_t wave_sum[kItemThreads / 64];
    const int per = (p.nbins + kItemThreads - 1) / kItemThreads;
    const int b0 = threadIdx.x * per;
    const int b1 = min(p.nbins, b0 + per);
    auto bin_start = [&](int b) -> uint32_t {
        return b < p.nbins ? p.offs[static_cast<int64_t>(b) * p.nblocks] : static_cast<uint32_t>(p.npts);
    };
    uint32_t mine = 0;
    for (int b = b0; b < b1; ++b) {
        const uint32_t n = bin_start(b + 1) - bin_start(b);
        mine += (n + kBrickItemPoints - 1) / kBrickItemPoints;
    }
    // exclusive scan of `mine` over the block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kItemThreads / 64; ++w) {
        if (w < wave) before += wave_sum[w];
        total += wave_sum[w];
    }
    uint32_t at = before + incl - mine;
    for (int b = b0; b < b1; ++b) {
        const uint32_t s = bin_start(b), n = bin_start(b + 1) - s;
        for (uint32_t q = 0; q < n; q += kBrickItemPoints) {
            if (at < p.max_items)
                p.items[at] = BrickItem{s + q, min(n - q, static_cast<uint32_t>(kBrickItemPoints)),
                                        static_cast<uint32_t>(b), 0u};
            ++at;
        }
    }
    if (threadIdx.x == 0) *p.item_count = min(total, static_cast<uint32_t>(p.max_items));
}

}  // namespace

// Blocks of the histogram and scatter passes: a few per CU, each over a
// fixed range of at least 4096 points, with the bin-major count table
// within one scan (nbins * nblocks <= 2^24).
int64_t group_blocks(int64_t npts, int nbins) {
    int64_t g = 1024;
    while (g > 1 && (g * nbins > (int64_t{1} << 24) || g * 4096 > npts)) g /= 2;
    return g;
}

int64_t group_max_items(int64_t npts, int nbins) { return (npts + kBrickItemPoints - 1) / kBrickItemPoints + nbins; }

hipError_t launch_grouping(GroupParams p, hipStream_t s) {
    if (p.nbins < 1 || p.nbins > kGroupMaxBins || p.nblocks < 1 || p.npts < 1 || p.npts > 0xFFFFFFFFll)
        return hipErrorInvalidValue;
    p.per_block = (p.npts + p.nblocks - 1) / p.nblocks;
    const size_t lds = static_cast<size_t>(p.nbins) * sizeof(uint32_t);
    hipLaunchKernelGGL(group_hist_kernel, dim3(static_cast<unsigned>(p.nblocks)), dim3(kGroupThreads), lds, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = exclusive_scan_u32(p.hist, p.offs, p.tile_sums, static_cast<int64_t>(p.nbins) * p.nblocks, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(group_scatter_kernel, dim3(static_cast<unsigned>(p.nblocks)), dim3(kGroupThreads), lds, s, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(group_items_kernel, dim3(1), dim3(kItemThreads), 0, s, p);
    return hipGetLastError();
}

}  // namespace intp_amd
