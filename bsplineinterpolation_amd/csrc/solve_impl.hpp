// solve_impl.hpp -- control-point solve on gfx950 (device code shared by
// solve.hip and the batched kernels of batch.hip): every grid line of one axis
// is an independent band (or cyclic band) system sharing one LU
// factorization (InterpolationTemplate.hpp:448-580).  One lane solves one
// line.
//
// Contiguous axis (solve_tile_kernel): a workgroup copies the lines of up
// to 64 neighbouring lanes into LDS with coalesced bulk loads (runs of one
// line), each lane substitutes its line in LDS, and the workgroup writes
// the tile back the same way.  The first pass reads the caller's values
// with the periodic index shift applied (InterpolationTemplate.hpp:452-462),
// so no separate copy precedes it.  Strided axes with 64 or more adjacent
// lines, and lines too long for LDS, use solve_kernel, which substitutes
// in place in HBM (coalesced across lanes on a strided axis).
//
// The substitution is the reference's (BandLU.hpp:121-143, 216-259)
// reordered by rows: each entry receives the same updates in the same
// order, the last `bw` values ride in registers, and the cyclic strips are
// register accumulators (forward) and register-held corner values
// (backward).  Built without FMA contraction and with a true division by
// the pivot, the result is bit-identical to the reference's.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace intp_amd {

namespace {

constexpr int kSolveBlock = 256;
constexpr int kTileThreads = 256;
constexpr int kMaxTileLines = 64;

// Elements a lane moves per batch: their loads (values and factor
// coefficients) are issued together, then substituted in order, then
// stored.
constexpr int kBatch = 8;

template <int BW, bool CYC, typename C>
struct LineState {
    C hist[BW + 1];    // forward: hist[m-1] = y[j-m]
    C acc[BW + 1];     // forward: rows n-BW+r of a cyclic matrix, strip-updated
    C nxt[BW + 1];     // backward: nxt[m-1] = x[j+m]
    C corner[BW + 1];  // backward: x[n-BW+c]
};

// A zero the compiler cannot see through, held in a vector register.  The
// factor coefficients are the same for every lane, so without it they
// would be read with scalar loads, whose completion counter is shared with
// LDS reads and not ordered: every batch would then wait for all loads in
// flight, including the next batch's.  Offset by this zero they are
// vector loads, which complete in order, so the next batch stays in
// flight while the current one is substituted.
__device__ __forceinline__ int32_t lane_zero() {
    int32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// Factor coefficients live in global memory.  Pointers read from a
// spline record (the batched kernels) are generic, and generic loads are
// flat loads, which count on the LDS counter as well: the same waits as
// scalar loads.  The explicit address space keeps them global loads.
template <typename C>
using GlobalPtr = const __attribute__((address_space(1))) C*;

template <typename C>
__device__ __forceinline__ GlobalPtr<C> factor_ptr(const void* p, int32_t z) {
    return (GlobalPtr<C>)(p) + z;
}

// Registers of one batch of forward rows: the loaded right-hand sides and
// the factor coefficients of those rows.
template <int BW, bool CYC, typename C, int B>
struct ForwardBatch {
    C xb[B];
    C lc[B][BW + 1];
    C sc[B][BW + 1];
};

// Registers of one batch of backward rows.
template <int BW, bool CYC, typename C, int B>
struct BackwardBatch {
    C xb[B];
    C uc[B][BW + 1];
    C rc[B][BW + 1];
    C dg[B];
};

template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void load_forward(ForwardBatch<BW, CYC, C, B>& r, const P* __restrict__ x, int64_t step,
                                             int64_t n, int64_t j0, const SolveParams& sp, int32_t z) {
    const GlobalPtr<C> __restrict__ lower = factor_ptr<C>(sp.lower, z);
    const GlobalPtr<C> __restrict__ bottom = factor_ptr<C>(sp.bottom, z);
#pragma unroll
    for (int u = 0; u < B; ++u) {
        r.xb[u] = static_cast<C>(x[(j0 + u) * step]);
#pragma unroll
        for (int m = 0; m < BW; ++m) r.lc[u][m] = lower[(j0 + u) * BW + m];
        if constexpr (CYC) {
#pragma unroll
            for (int q = 0; q < BW; ++q) r.sc[u][q] = bottom[q * n + j0 + u];
        }
    }
}

// Forward substitution of rows j0 .. j0+B-1 (unit lower triangle, plus
// the bottom strip of a cyclic matrix) from loaded registers; stores the
// rows.
template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void run_forward(ForwardBatch<BW, CYC, C, B>& r, P* __restrict__ x, int64_t step,
                                            int64_t n, int64_t j0, LineState<BW, CYC, C>& st) {
#pragma clang fp contract(off)
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 + u;
        C v = r.xb[u];
        if constexpr (CYC) {
#pragma unroll
            for (int q = 0; q < BW; ++q)
                if (j == n - BW + q) v = st.acc[q];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - r.lc[u][m - 1] * st.hist[m - 1];
        r.xb[u] = v;
#pragma unroll
        for (int m = BW; m >= 2; --m) st.hist[m - 1] = st.hist[m - 2];
        if (BW > 0) st.hist[0] = v;
        if constexpr (CYC) {
#pragma unroll
            for (int q = 0; q < BW; ++q)
                if (j < n - BW + q - BW) st.acc[q] = st.acc[q] - r.sc[u][q] * v;
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) x[(j0 + u) * step] = static_cast<P>(r.xb[u]);
}

template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void load_backward(BackwardBatch<BW, CYC, C, B>& r, const P* __restrict__ x, int64_t step,
                                              int64_t n, int64_t j0, const SolveParams& sp, int32_t z) {
    const GlobalPtr<C> __restrict__ upper = factor_ptr<C>(sp.upper, z);
    const GlobalPtr<C> __restrict__ diag = factor_ptr<C>(sp.diag, z);
    const GlobalPtr<C> __restrict__ right = factor_ptr<C>(sp.right, z);
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 - u;
        r.xb[u] = static_cast<C>(x[j * step]);
#pragma unroll
        for (int m = 0; m < BW; ++m) r.uc[u][m] = upper[j * BW + m];
        r.dg[u] = diag[j];
        if constexpr (CYC) {
#pragma unroll
            for (int c = 0; c < BW; ++c) r.rc[u][c] = right[c * n + j];
        }
    }
}

// Backward substitution of rows j0, j0-1, .., j0-B+1 (upper triangle,
// plus the right strip of a cyclic matrix) from loaded registers; stores
// the rows.
template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void run_backward(BackwardBatch<BW, CYC, C, B>& r, P* __restrict__ x, int64_t step,
                                             int64_t n, int64_t j0, LineState<BW, CYC, C>& st) {
#pragma clang fp contract(off)
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int64_t j = j0 - u;
        C v = r.xb[u];
        if constexpr (CYC) {
#pragma unroll
            for (int c = BW - 1; c >= 0; --c)
                if (j < n - BW + c - BW) v = v - r.rc[u][c] * st.corner[c];
        }
#pragma unroll
        for (int m = BW; m >= 1; --m) v = v - r.uc[u][m - 1] * st.nxt[m - 1];
        v = v / r.dg[u];
        r.xb[u] = v;
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
    for (int u = 0; u < B; ++u) x[(j0 - u) * step] = static_cast<P>(r.xb[u]);
}

// Rows per batch: two batches are in registers at once (the one being
// substituted and the next one loading), so wide cyclic bands use 4.
// Kernels that merge many band widths pass 4 to bound their registers.
template <int BW, bool CYC>
constexpr int batch_rows() {
    return BW * (CYC ? 2 : 1) <= 4 ? kBatch : kBatch / 2;
}

// Forward substitution of rows a .. e-1 of a line from state `st`
// (software-pipelined: the loads of the next batch of rows, right-hand
// sides and factor coefficients, are in flight while the current batch is
// substituted, so one lane's chain of dependent rows does not wait for
// memory every batch; each row receives the same updates in the same
// order as without pipelining).
template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void forward_range(P* __restrict__ x, int64_t step, int64_t n, int64_t a, int64_t e,
                                              const SolveParams& sp, LineState<BW, CYC, C>& st, int32_t z) {
    const int64_t nb = (e - a) / B;  // full batches
    if (nb > 0) {
        ForwardBatch<BW, CYC, C, B> ba, bb;
        load_forward<BW, CYC, C, P, B>(ba, x, step, n, a, sp, z);
        int64_t t = 0;
        for (; t + 2 <= nb; t += 2) {
            load_forward<BW, CYC, C, P, B>(bb, x, step, n, a + (t + 1) * B, sp, z);
            run_forward<BW, CYC, C, P, B>(ba, x, step, n, a + t * B, st);
            // batch t + 2, or (past the end) a harmless reload of batch nb - 1
            load_forward<BW, CYC, C, P, B>(ba, x, step, n, a + min(t + 2, nb - 1) * B, sp, z);
            run_forward<BW, CYC, C, P, B>(bb, x, step, n, a + (t + 1) * B, st);
        }
        if (t < nb) run_forward<BW, CYC, C, P, B>(ba, x, step, n, a + t * B, st);
    }
    for (int64_t j = a + nb * B; j < e; ++j) {
        ForwardBatch<BW, CYC, C, 1> r;
        load_forward<BW, CYC, C, P, 1>(r, x, step, n, j, sp, z);
        run_forward<BW, CYC, C, P, 1>(r, x, step, n, j, st);
    }
}

// Backward substitution of rows e-1 down to a from state `st` (pipelined
// like forward_range).
template <int BW, bool CYC, typename C, typename P, int B>
__device__ __forceinline__ void backward_range(P* __restrict__ x, int64_t step, int64_t n, int64_t a, int64_t e,
                                               const SolveParams& sp, LineState<BW, CYC, C>& st, int32_t z) {
    const int64_t nb = (e - a) / B;
    if (nb > 0) {
        BackwardBatch<BW, CYC, C, B> ba, bb;
        load_backward<BW, CYC, C, P, B>(ba, x, step, n, e - 1, sp, z);
        int64_t t = 0;
        for (; t + 2 <= nb; t += 2) {
            load_backward<BW, CYC, C, P, B>(bb, x, step, n, e - 1 - (t + 1) * B, sp, z);
            run_backward<BW, CYC, C, P, B>(ba, x, step, n, e - 1 - t * B, st);
            load_backward<BW, CYC, C, P, B>(ba, x, step, n, e - 1 - min(t + 2, nb - 1) * B, sp, z);
            run_backward<BW, CYC, C, P, B>(bb, x, step, n, e - 1 - (t + 1) * B, st);
        }
        if (t < nb) run_backward<BW, CYC, C, P, B>(ba, x, step, n, e - 1 - t * B, st);
    }
    for (int64_t j = e - 1 - nb * B; j >= a; --j) {
        BackwardBatch<BW, CYC, C, 1> r;
        load_backward<BW, CYC, C, P, 1>(r, x, step, n, j, sp, z);
        run_backward<BW, CYC, C, P, 1>(r, x, step, n, j, st);
    }
}

// Solves one line x[0], x[step], ..., x[(n-1)*step] in place, computing
// in C (elements of x have type P).
template <int BW, bool CYC, typename C, typename P, int B = batch_rows<BW, CYC>()>
__device__ __forceinline__ void substitute(P* __restrict__ x, int64_t step, int64_t n, const SolveParams& sp) {
    LineState<BW, CYC, C> st;
#pragma unroll
    for (int m = 0; m <= BW; ++m) st.hist[m] = st.nxt[m] = st.corner[m] = st.acc[m] = C(0);
    if constexpr (CYC) {
#pragma unroll
        for (int q = 0; q < BW; ++q) st.acc[q] = static_cast<C>(x[(n - BW + q) * step]);
    }
    const int32_t z = lane_zero();
    forward_range<BW, CYC, C, P, B>(x, step, n, 0, n, sp, st, z);
    backward_range<BW, CYC, C, P, B>(x, step, n, 0, n, sp, st, z);
}

// A launch of a few long lines (a long 1D spline, say) substitutes each
// line with one lane whose factor loads run one batch of rows ahead: from
// HBM every batch would wait a full memory latency.  When the launch has
// at most kWarmLines lines and the factors fit kWarmBytes, every thread of
// the workgroup first reads the factor arrays once (one 4-byte word per
// 64 bytes), so that the substitution finds them in L2.  The arithmetic is
// unchanged.
constexpr int64_t kWarmLines = 8;
// 32 MB: past the 4 MB L2 of an XCD the factors still land in the
// device's last-level cache (1D n = 100,000 cubic: 12.8 -> 12.0 ms
// against a 2 MB limit; n <= 20,000 unchanged).
constexpr int64_t kWarmBytes = int64_t{32} << 20;

template <int BW, bool CYC, typename C, int NT>
__device__ __forceinline__ void warm_factors(const SolveParams& sp, int64_t n) {
    if (sp.nlines > kWarmLines) return;
    const int64_t per = static_cast<int64_t>(sizeof(C)) * n;
    const int64_t bytes = per * (2 * BW + 1) + (CYC ? per * 2 * BW : 0);
    if (bytes > kWarmBytes) return;
    uint32_t acc = 0;
    auto touch = [&](const void* p, int64_t nbytes) {
        const uint32_t* w = static_cast<const uint32_t*>(p);
        const int64_t lines = nbytes / 64;
        for (int64_t i = threadIdx.x; i < lines; i += NT * 4) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = i + u * NT < lines ? w[(i + u * NT) * 16] : 0u;
#pragma unroll
            for (int u = 0; u < 4; ++u) acc ^= v[u];
        }
    };
    touch(sp.lower, per * BW);
    touch(sp.upper, per * BW);
    touch(sp.diag, per);
    if constexpr (CYC) {
        touch(sp.bottom, per * BW);
        touch(sp.right, per * BW);
    }
    asm volatile("" ::"v"(acc));  // the loads' results are used
}

// Fallback: one lane per line, substituting in HBM (storage T, compute C;
// with T narrower than C the forward results are rounded to T between
// the two sweeps).

// Odd LDS row stride (in elements), so that lanes reading the same column
// of their rows hit distinct banks.
__host__ __device__ __forceinline__ int64_t tile_stride(int64_t n) { return n | 1; }

// One workgroup: lines [blockIdx.x * lpb, + lpb) staged through LDS in
// the compute type C; the table holds T.  With sp.accum set, the solved
// lines are added into accum instead of being stored to x.
// WARM: warm_factors before the substitution (single-spline launches).
template <int BW, bool CYC, typename T, typename C, int B = batch_rows<BW, CYC>(), int NT = kTileThreads,
          bool WARM = false>
__device__ __forceinline__ void solve_tile(const SolveParams& sp, int64_t first, unsigned char* tile_bytes,
                                           int64_t* dst_base, int64_t* src_base) {
    C* tile = reinterpret_cast<C*>(tile_bytes);
    T* __restrict__ X = static_cast<T*>(sp.x);
    const int lpb = sp.lines_per_block;
    const int64_t n = sp.n, inner = sp.inner, stride = tile_stride(n);
    const int nl = static_cast<int>(min(static_cast<int64_t>(lpb), sp.nlines - first));
    const int tid = threadIdx.x;
    if (tid < nl) {
        const int64_t line = first + tid;
        const int64_t o = line / inner, ii = line - o * inner;
        dst_base[tid] = o * n * inner + ii;
        if (sp.src) {
            // inner == 1: o runs over the outer dims in C order
            // constant indices (a runtime-indexed loop would force a
            // SolveParams built in registers out to scratch memory)
            int64_t rem = o, off = 0, mul = n;
#pragma unroll
            for (int d = kMaxDim - 1; d >= 0; --d) {
                if (d >= sp.outer_dims) continue;
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
    // Strided lines: (line e % lpb, element e / lpb); as NT % lpb == 0 a
    // thread keeps one line.
    const uint32_t un = static_cast<uint32_t>(n);
    const int lshift = __builtin_ctz(static_cast<unsigned>(lpb));
    const uint32_t count = static_cast<uint32_t>(inner == 1 ? nl * n : lpb * n);
    const uint32_t my_l = static_cast<uint32_t>(tid) & static_cast<uint32_t>(lpb - 1);
    if (sp.src) {
        for (uint32_t e0 = tid; e0 < count; e0 += NT * kBatch) {
            int64_t from[kBatch];
            uint32_t at[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = min(e0 + u * NT, count - 1);
                const uint32_t l = e / un, jj = e - l * un;
                int64_t js = static_cast<int64_t>(jj) - sp.shift_last;
                if (js < 0) js += n;
                from[u] = src_base[l] + js;
                at[u] = l * static_cast<uint32_t>(stride) + jj;
            }
            // the values are read once: non-temporal loads keep them from
            // evicting the factors and the lines being written
            C v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = static_cast<C>(__builtin_nontemporal_load(&sp.src[from[u]]));
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (e0 + u * NT < count) tile[at[u]] = v[u];
        }
    } else if (inner == 1) {
        const T* __restrict__ base = X + first * n;
        for (uint32_t e0 = tid; e0 < count; e0 += NT * kBatch) {
            C v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = static_cast<C>(base[min(e0 + u * NT, count - 1)]);
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * NT;
                if (e < count) {
                    const uint32_t l = e / un, jj = e - l * un;
                    tile[l * static_cast<uint32_t>(stride) + jj] = v[u];
                }
            }
        }
    } else if (static_cast<int>(my_l) < nl) {
        const T* __restrict__ line = X + dst_base[my_l];
        C* __restrict__ row = tile + my_l * stride;
        for (uint32_t e0 = tid; e0 < count; e0 += NT * kBatch) {
            C v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t jj = min(e0 + u * NT, count - 1) >> lshift;
                v[u] = static_cast<C>(line[static_cast<int64_t>(jj) * inner]);
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * NT;
                if (e < count) row[e >> lshift] = v[u];
            }
        }
    }
    if constexpr (WARM) warm_factors<BW, CYC, C, NT>(sp, n);
    __syncthreads();

    if (tid < nl) substitute<BW, CYC, C, C, B>(tile + tid * stride, 1, n, sp);
    __syncthreads();

    // ---- store ----
    T* __restrict__ Y = sp.accum ? static_cast<T*>(sp.accum) : X;
    const bool add = sp.accum != nullptr;
    if (inner == 1) {
        T* __restrict__ base = Y + first * n;
        for (uint32_t e0 = tid; e0 < count; e0 += NT * kBatch) {
            C v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = min(e0 + u * NT, count - 1);
                const uint32_t l = e / un, jj = e - l * un;
                v[u] = tile[l * static_cast<uint32_t>(stride) + jj];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (e0 + u * NT < count) {
                    T* at = base + e0 + u * NT;
                    *at = add ? static_cast<T>(*at + static_cast<T>(v[u])) : static_cast<T>(v[u]);
                }
        }
    } else if (static_cast<int>(my_l) < nl) {
        T* __restrict__ line = Y + dst_base[my_l];
        const C* __restrict__ row = tile + my_l * stride;
        for (uint32_t e0 = tid; e0 < count; e0 += NT * kBatch) {
            C v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) v[u] = row[min(e0 + u * NT, count - 1) >> lshift];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t e = e0 + u * NT;
                if (e < count) {
                    T* at = line + static_cast<int64_t>(e >> lshift) * inner;
                    *at = add ? static_cast<T>(*at + static_cast<T>(v[u])) : static_cast<T>(v[u]);
                }
            }
        }
    }
}

}  // namespace

}  // namespace intp_amd
