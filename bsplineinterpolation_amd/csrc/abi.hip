// abi.hip -- the C ABI of libintp_amd.so (declared in include/intp_amd.h):
// handles, device memory, argument checks and kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/intp_amd.h"
#include "kernels.hpp"
#include "grid.hpp"
#include "host_eval.hpp"
#include "runtime.hpp"
#include "setup.hpp"

using namespace intp_amd;

namespace intp_amd {

hipError_t launch_eval(int dim, int order, bool grad, bool uniform, bool f32, const EvalParams& p, hipStream_t s) {
    switch (dim * 2 + (f32 ? 1 : 0)) {
        case 2: return eval_dispatch_f64_d1(order, grad, uniform, p, s);
        case 3: return eval_dispatch_f32_d1(order, grad, uniform, p, s);
        case 4: return eval_dispatch_f64_d2(order, grad, uniform, p, s);
        case 5: return eval_dispatch_f32_d2(order, grad, uniform, p, s);
        case 6: return eval_dispatch_f64_d3(order, grad, uniform, p, s);
        case 7: return eval_dispatch_f32_d3(order, grad, uniform, p, s);
        case 8: return eval_dispatch_f64_d4(order, grad, uniform, p, s);
        case 9: return eval_dispatch_f32_d4(order, grad, uniform, p, s);
        case 10: return eval_dispatch_f64_d5(order, grad, uniform, p, s);
        case 11: return eval_dispatch_f32_d5(order, grad, uniform, p, s);
        case 12: return eval_dispatch_f64_d6(order, grad, uniform, p, s);
        case 13: return eval_dispatch_f32_d6(order, grad, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_eval_grad(int dim, int order, bool uniform, bool f32, const EvalParams& p, hipStream_t s) {
    return launch_eval(dim, order, true, uniform, f32, p, s);
}

hipError_t launch_eval_single(int dim, int order, bool uniform, bool f32, const EvalParams& p, hipStream_t s) {
    return launch_eval(dim, order, false, uniform, f32, p, s);
}

// value and first partials of a refined build from its gradient table
hipError_t launch_eval_cells(int dim, int order, bool uniform, const EvalParams& p, hipStream_t s) {
    switch (dim) {
        case 1: return eval_cells_d1(order, uniform, p, s);
        case 2: return eval_cells_d2(order, uniform, p, s);
        case 3: return eval_cells_d3(order, uniform, p, s);
        default: return hipErrorNotSupported;
    }
}

}  // namespace intp_amd

namespace {

// Knots, factors, reciprocal tables and basis polynomials of one grid on
// one device, packed into one buffer; shared by a template and every
// spline interpolated from it.
struct GridData {
    int device = 0;
    int dim = 0, order = 0;
    DimSetup setup[kMaxDim];
    // views into `dev` (nullptr where the table is empty)
    const double* knots[kMaxDim] = {};
    const double* recip[kMaxDim] = {};
    const double* lower[kMaxDim] = {};
    const double* upper[kMaxDim] = {};
    const double* diag[kMaxDim] = {};
    const double* bottom[kMaxDim] = {};
    const double* right[kMaxDim] = {};
    const double* poly[kMaxDim] = {};
    const double* colloc[kMaxDim] = {};  // unfactored collocation rows (refined builds)
    const float* lower32[kMaxDim] = {};   // LU factors rounded to FP32
    const float* upper32[kMaxDim] = {};
    const float* diag32[kMaxDim] = {};
    const float* bottom32[kMaxDim] = {};
    const float* right32[kMaxDim] = {};
    bool has_factors = false;
    int64_t total = 1;
    std::vector<unsigned char> image;  // host copy of `dev`, the source of its upload
    unsigned char* dev = nullptr;
    hipEvent_t uploaded = nullptr;
    UseTracker uses;

    ~GridData() {
        release_after(device, dev, uses);
        if (uploaded) {
            // the upload reads `image`, which is freed after this body
            DeviceScope scope(device);
            (void)hipEventSynchronize(uploaded);
            (void)hipEventDestroy(uploaded);
        }
    }
};

// Packs every table of `g` (each 256-byte aligned) into one image and
// uploads it with one asynchronous copy on `stream`.
int upload_grid(GridData& g, hipStream_t stream) {
    struct Item {
        const void* src;
        size_t bytes;
        std::function<void(const unsigned char*)> set_view;
    };
    std::vector<Item> items;
    auto add = [&](const auto& v, auto& view) {
        using P = std::remove_reference_t<decltype(view)>;
        items.push_back({v.data(), v.size() * sizeof(v[0]),
                         [&view](const unsigned char* at) { view = reinterpret_cast<P>(at); }});
    };
    for (int d = 0; d < g.dim; ++d) {
        DimSetup& s = g.setup[d];
        add(s.knots, g.knots[d]);
        add(s.recip, g.recip[d]);
        add(s.poly, g.poly[d]);
        if (g.has_factors) {
            add(s.lu.lower, g.lower[d]);
            add(s.lu.upper, g.upper[d]);
            add(s.lu.diag, g.diag[d]);
            add(s.lu.bottom, g.bottom[d]);
            add(s.lu.right, g.right[d]);
            add(s.colloc, g.colloc[d]);
            add(s.lower32, g.lower32[d]);
            add(s.upper32, g.upper32[d]);
            add(s.diag32, g.diag32[d]);
            add(s.bottom32, g.bottom32[d]);
            add(s.right32, g.right32[d]);
        }
    }
    std::vector<size_t> offset;
    size_t total = 0;
    for (const Item& it : items) {
        offset.push_back(total);
        total += (it.bytes + 255) / 256 * 256;
    }
    g.image.assign(total, 0);
    for (size_t i = 0; i < items.size(); ++i)
        if (items[i].bytes) std::memcpy(g.image.data() + offset[i], items[i].src, items[i].bytes);
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&g.dev), total, stream), "device allocation (grid tables)");
    HIP_TRY(hipMemcpyAsync(g.dev, g.image.data(), total, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    HIP_TRY(g.uses.record(stream), "hipEventRecord");
    HIP_TRY(hipEventCreateWithFlags(&g.uploaded, hipEventDisableTiming), "hipEventCreate");
    HIP_TRY(hipEventRecord(g.uploaded, stream), "hipEventRecord");
    for (size_t i = 0; i < items.size(); ++i) items[i].set_view(items[i].bytes ? g.dev + offset[i] : nullptr);
    return INTP_OK;
}

// ---- INTP_EVAL_TIMED: events around the grouping and the kernels ---------

struct TimedCall {
    int device;
    hipEvent_t ev[3];  // before grouping, before the eval kernels, after them
};

thread_local std::vector<TimedCall> t_timed;
thread_local std::vector<std::pair<int, hipEvent_t>> t_spare_events;

hipError_t timing_event(int device, hipEvent_t* out) {
    for (size_t i = 0; i < t_spare_events.size(); ++i)
        if (t_spare_events[i].first == device) {
            *out = t_spare_events[i].second;
            t_spare_events.erase(t_spare_events.begin() + static_cast<std::ptrdiff_t>(i));
            return hipSuccess;
        }
    return hipEventCreate(out);
}

}  // namespace


struct intp_template_s {
    std::shared_ptr<GridData> grid;
};

struct intp_spline_s {
    std::atomic<int> refs{1};
    std::shared_ptr<GridData> grid;
    void* ctrl = nullptr;  // FP64, or FP32 for a refined build
    // refined builds: the gradient table (Cell1 / Cell3 per control
    // point), in the allocation of `ctrl` after the FP32 control points
    void* cells = nullptr;
    int ctrl_bytes = 8;
    int sweeps = -1;       // refinement sweeps of a refined build
    UseTracker uses;       // streams that read `ctrl` (solve and every eval)
    // Host copy of the control points (FP64), downloaded on the first
    // host-side evaluation; immutable afterwards like the device table.
    std::once_flag mirror_once;
    std::atomic<bool> mirror_ready{false};  // set once `mirror` holds the table
    int mirror_rc = INTP_OK;
    std::string mirror_error;
    std::vector<double> mirror;
    // refined splines with a gradient table: its differences per dimension
    // (mirror_diff[d * total + e]), downloaded with the mirror
    std::vector<double> mirror_diff;
    ~intp_spline_s() {
        // ordered on the release stream before the grid's own release,
        // which runs when `grid` drops its last reference below
        release_after(grid->device, ctrl, uses);
    }
};

namespace {

int check_dims(int dim, int order) {
    if (dim < 1 || dim > kMaxDim)
        return fail(INTP_E_UNSUPPORTED, "dimension " + std::to_string(dim) + " has no compiled kernels (1.." +
                                            std::to_string(kMaxDim) + ")");
    if (order < 0 || order > kMaxOrder)
        return fail(INTP_E_UNSUPPORTED, "order " + std::to_string(order) + " has no compiled kernels (0.." +
                                            std::to_string(kMaxOrder) + ")");
    return INTP_OK;
}

// Knots, factored collocation matrices and basis tables of `desc`,
// uploaded on `stream` (on the device's setup stream, waited for, when
// `stream` is null: a template outlives any one stream).
int make_grid(intp_ctx ctx, const intp_grid_desc* desc, hipStream_t stream, bool own_stream,
              std::shared_ptr<GridData>& out) {
    if (!ctx || !desc) return fail(INTP_E_INVALID, "null argument");
    int rc = check_dims(desc->dim, desc->order);
    if (rc != INTP_OK) return rc;
    auto g = std::make_shared<GridData>();
    g->device = ctx->device;
    g->dim = desc->dim;
    g->order = desc->order;
    g->has_factors = true;
    for (int d = 0; d < desc->dim; ++d) {
        const std::string err = setup_dimension(desc->order, desc->periodic[d] != 0, desc->n[d], desc->xmin[d],
                                                desc->xmax[d], desc->coords[d], g->setup[d]);
        if (!err.empty()) return fail(INTP_E_INVALID, err + " (dimension " + std::to_string(d) + ")");
        g->total *= desc->n[d];
    }
    DeviceScope scope(ctx->device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    DeviceRuntime* rt = device_runtime(ctx->device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    if (own_stream) stream = rt->setup;
    if ((rc = upload_grid(*g, stream)) != INTP_OK) return rc;
    if (own_stream) HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    out = std::move(g);
    return INTP_OK;
}

// ---- grids of recent builds ---------------------------------------------
//
// intp_build / intp_build_refined on a uniform grid equal to a recent
// build's (same device, dimension, order, sizes, periodicity and ranges)
// reuse that grid's uploaded tables instead of recomputing and uploading
// them: the tables are a deterministic function of the descriptor, so the
// results do not change.  A few small grids are kept (grids whose tables
// exceed kGridCacheBytes are never cached); a spline keeps its grid alive
// whether or not the cache still holds it.
constexpr size_t kGridCacheEntries = 16;
constexpr size_t kGridCacheBytes = size_t{8} << 20;

struct GridKey {
    int device = 0, dim = 0, order = 0;
    int64_t n[kMaxDim] = {};
    int32_t periodic[kMaxDim] = {};
    double xmin[kMaxDim] = {}, xmax[kMaxDim] = {};
    bool operator==(const GridKey& o) const {
        if (device != o.device || dim != o.dim || order != o.order) return false;
        for (int d = 0; d < dim; ++d)
            if (n[d] != o.n[d] || periodic[d] != o.periodic[d] || std::memcmp(&xmin[d], &o.xmin[d], 8) != 0 ||
                std::memcmp(&xmax[d], &o.xmax[d], 8) != 0)
                return false;
        return true;
    }
};

// Never destroyed: its grids would otherwise be freed by static
// destructors, after the HIP runtime may have shut down.
std::mutex g_grid_cache_mu;
std::list<std::pair<GridKey, std::shared_ptr<GridData>>>& g_grid_cache =
    *new std::list<std::pair<GridKey, std::shared_ptr<GridData>>>();  // front: most recent

// The key of a uniform descriptor; false for non-uniform grids (not cached).
bool grid_key(int device, const intp_grid_desc& desc, GridKey* key) {
    key->device = device;
    key->dim = desc.dim;
    key->order = desc.order;
    for (int d = 0; d < desc.dim; ++d) {
        if (desc.coords[d]) return false;
        key->n[d] = desc.n[d];
        key->periodic[d] = desc.periodic[d] != 0;
        key->xmin[d] = desc.xmin[d];
        key->xmax[d] = desc.xmax[d];
    }
    return true;
}

// A cached grid for `key`, made usable on `stream` (which waits for its
// upload), or null.
std::shared_ptr<GridData> cached_grid(const GridKey& key, hipStream_t stream) {
    std::shared_ptr<GridData> g;
    {
        std::lock_guard<std::mutex> lk(g_grid_cache_mu);
        for (auto it = g_grid_cache.begin(); it != g_grid_cache.end(); ++it)
            if (it->first == key) {
                g = it->second;
                g_grid_cache.splice(g_grid_cache.begin(), g_grid_cache, it);
                break;
            }
    }
    if (g && g->uploaded && hipStreamWaitEvent(stream, g->uploaded, 0) != hipSuccess) return nullptr;
    return g;
}

void remember_grid(const GridKey& key, const std::shared_ptr<GridData>& g) {
    if (g->image.size() > kGridCacheBytes) return;
    std::shared_ptr<GridData> evicted;  // released outside the lock
    std::lock_guard<std::mutex> lk(g_grid_cache_mu);
    g_grid_cache.emplace_front(key, g);
    if (g_grid_cache.size() > kGridCacheEntries) {
        evicted = std::move(g_grid_cache.back().second);
        g_grid_cache.pop_back();
    }
}

// make_grid for a build: a cached grid when the descriptor has one.
int build_grid(intp_ctx ctx, const intp_grid_desc* desc, hipStream_t stream, std::shared_ptr<GridData>& out) {
    GridKey key;
    const bool cacheable = ctx && desc && desc->dim >= 1 && desc->dim <= kMaxDim && grid_key(ctx->device, *desc, &key);
    if (cacheable) {
        DeviceScope scope(ctx->device);
        if (scope.ok() && (out = cached_grid(key, stream))) return INTP_OK;
    }
    const int rc = make_grid(ctx, desc, stream, false, out);
    if (rc == INTP_OK && cacheable) remember_grid(key, out);
    return rc;
}

// Periodic index shift of the right-hand sides along each dimension
// (InterpolationTemplate.hpp:452-462).
void solve_shifts(const GridData& g, int64_t (&shift)[kMaxDim]) {
    for (int d = 0; d < kMaxDim; ++d) shift[d] = 0;
    for (int d = 0; d < g.dim; ++d) shift[d] = g.setup[d].periodic ? (g.order / 2) % g.setup[d].n : 0;
}

// Element types of the solve along axis d: FP64 builds are FP64 through;
// refined builds store FP32 and compute in FP32 with FP32 factors, except
// periodic dimensions of order >= 5, whose factors stay FP64.
SolveTypes axis_types(const GridData& g, bool refined, int d) {
    if (!refined) return SolveTypes::kF64;
    return (g.setup[d].periodic && g.order >= 5) ? SolveTypes::kF32StoreF64Compute : SolveTypes::kF32;
}

int compute_bytes(SolveTypes t) { return t == SolveTypes::kF32 ? 4 : 8; }

// Line solves along every axis, last axis first (InterpolationTemplate.hpp:448-580),
// on table x (FP64, or FP32 when `refined`).  With `values` set, the first
// pass reads them (FP64) with the periodic shift applied; when its lines do
// not fit the LDS-tiled kernel, a shift copy precedes an in-place first
// pass.  Without `values`, x already holds shifted right-hand sides.  With
// `accum` set, the last pass adds its result into accum instead of x.
int solve_axes(const GridData& g, bool refined, const double* values, void* x, void* accum, hipStream_t stream) {
    const int last = g.dim - 1;
    int64_t shift[kMaxDim];
    solve_shifts(g, shift);
    bool any_shift = false;
    for (int d = 0; d < g.dim; ++d) any_shift = any_shift || shift[d] != 0;
    const size_t elem = refined ? sizeof(float) : sizeof(double);
    const bool fused =
        values != nullptr && solve_tile_lines(g.setup[last].n, compute_bytes(axis_types(g, refined, last))) > 0;
    if (values != nullptr && !fused) {
        if (any_shift || refined) {
            ShiftParams sh{};
            sh.src = values;
            sh.dst = x;
            sh.total = g.total;
            sh.dim = g.dim;
            for (int d = 0; d < g.dim; ++d) {
                sh.n[d] = g.setup[d].n;
                sh.shift[d] = shift[d];
            }
            HIP_TRY(launch_shift_copy(sh, refined, stream), "shift copy launch");
        } else if (values != x) {
            HIP_TRY(hipMemcpyAsync(x, values, static_cast<size_t>(g.total) * elem, hipMemcpyDeviceToDevice, stream),
                    "hipMemcpyAsync");
        }
    }
    for (int a = last; a >= 0; --a) {
        const DimSetup& s = g.setup[a];
        const SolveTypes types = axis_types(g, refined, a);
        const bool f32 = types == SolveTypes::kF32;
        SolveParams sp{};
        sp.x = x;
        sp.n = s.n;
        int64_t inner = 1, outer = 1;
        for (int d = a + 1; d < g.dim; ++d) inner *= g.setup[d].n;
        for (int d = 0; d < a; ++d) outer *= g.setup[d].n;
        sp.inner = inner;
        sp.nlines = inner * outer;
        sp.lower = f32 ? static_cast<const void*>(g.lower32[a]) : g.lower[a];
        sp.upper = f32 ? static_cast<const void*>(g.upper32[a]) : g.upper[a];
        sp.diag = f32 ? static_cast<const void*>(g.diag32[a]) : g.diag[a];
        sp.bottom = f32 ? static_cast<const void*>(g.bottom32[a]) : g.bottom[a];
        sp.right = f32 ? static_cast<const void*>(g.right32[a]) : g.right[a];
        sp.lines_per_block = solve_tile_lines(s.n, compute_bytes(types));
        // Strided axes with a wavefront or more of adjacent lines: lane l's
        // line sits next to lane l+1's, so substituting in place reads and
        // writes coalesced rows, and every wave of the workgroup substitutes
        // (the LDS tile keeps one wave busy while the other three wait).
        // The accumulating last pass keeps the tile, which adds on store.
        if (inner >= 64 && !(a == last && fused) && !(a == 0 && accum != nullptr)) sp.lines_per_block = 0;
        if (a == last && fused) {
            sp.src = values;
            sp.outer_dims = last;
            for (int d = 0; d < last; ++d) {
                sp.outer_n[d] = g.setup[d].n;
                sp.outer_shift[d] = shift[d];
            }
            sp.shift_last = shift[last];
        }
        const bool add_after = a == 0 && accum != nullptr && sp.lines_per_block == 0;
        if (a == 0 && accum != nullptr && !add_after) sp.accum = accum;
        hipError_t e = launch_solve(s.lu.p, s.lu.cyclic, types, sp, stream);
        if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "band width without a compiled solver");
        HIP_TRY(e, "solve launch");
        if (add_after) {
            if (!refined) return fail(INTP_E_INVALID, "accumulation is defined for FP32 tables only");
            HIP_TRY(launch_add_f32(static_cast<float*>(accum), static_cast<const float*>(x), g.total, stream),
                    "correction launch");
        }
    }
    return INTP_OK;
}

// FP64 control-point solve of `values` into `ctrl` (both device).
int solve_into(const GridData& g, const double* values, double* ctrl, hipStream_t stream) {
    return solve_axes(g, false, values, ctrl, nullptr, stream);
}

// Refined-FP32 control points of `values` into `ctrl` (FP32): an FP32
// solve, then `sweeps` rounds of an FP64 residual r = S f - (M_0 (x) ...) c
// and an FP32 correction solve added to c; and the gradient table `cells`.
// The last round's correction d is kept apart: the cells take their
// differences from c + d in FP64 before c becomes c + d rounded.
// `scratch` holds N floats and 2 N doubles (N doubles for 1D).
int refine_into(const GridData& g, const double* values, float* ctrl, void* cells, int sweeps,
                unsigned char* scratch, hipStream_t stream) {
    int rc = solve_axes(g, true, values, ctrl, nullptr, stream);
    if (rc != INTP_OK) return rc;
    CellParams cp{};
    cp.c0 = ctrl;
    cp.cells = cells;
    cp.dim = g.dim;
    cp.total = g.total;
    for (int d = 0; d < g.dim; ++d) {
        cp.n[d] = g.setup[d].n;
        cp.periodic[d] = g.setup[d].periodic ? 1 : 0;
    }
    float* delta = reinterpret_cast<float*>(scratch);
    double* t[2] = {reinterpret_cast<double*>(scratch + ((g.total * 4 + 255) / 256) * 256), nullptr};
    t[1] = t[0] + ((g.total + 31) / 32) * 32;
    int64_t shift[kMaxDim];
    solve_shifts(g, shift);
    for (int sweep = 0; sweep < sweeps; ++sweep) {
        // residual, axes last to first: c -> t -> ... -> delta
        const void* in = ctrl;
        bool in_f32 = true;
        for (int a = g.dim - 1; a >= 0; --a) {
            MatvecParams mp{};
            const DimSetup& s = g.setup[a];
            mp.x = in;
            mp.colloc = g.colloc[a];
            mp.w = s.colloc_w;
            mp.periodic = s.periodic ? 1 : 0;
            mp.n = s.n;
            int64_t inner = 1;
            for (int d = a + 1; d < g.dim; ++d) inner *= g.setup[d].n;
            mp.inner = inner;
            mp.total = g.total;
            const bool final_pass = a == 0;
            if (final_pass) {
                mp.y = delta;
                mp.src = values;
                mp.dim = g.dim;
                for (int d = 0; d < g.dim; ++d) {
                    mp.dn[d] = g.setup[d].n;
                    mp.dshift[d] = shift[d];
                }
            } else {
                mp.y = t[a % 2];
            }
            HIP_TRY(launch_band_matvec(mp, in_f32, final_pass, stream), "residual launch");
            in = mp.y;
            in_f32 = false;
        }
        // correction: solve on delta (already shifted), added into ctrl
        // (the last one by the cells' launch)
        const bool keep = sweep + 1 == sweeps && cells != nullptr;
        if ((rc = solve_axes(g, true, nullptr, delta, keep ? nullptr : ctrl, stream)) != INTP_OK) return rc;
        if (keep) cp.delta = delta;
    }
    if (cells) HIP_TRY(launch_refined_cells(cp, stream), "gradient table launch");
    return INTP_OK;
}

DimParams dim_params(const GridData& g, int d) {
    int64_t stride = 1;
    for (int e = d + 1; e < g.dim; ++e) stride *= g.setup[e].n;
    return make_dim_params(g.setup[d], g.order, stride, g.knots[d], g.recip[d], g.poly[d]);
}

// Batches at least this large are grouped by brick before evaluation when
// the table is large enough for grouping to pay (group_pays).
constexpr int64_t kGroupMinPoints = int64_t{1} << 20;

// Whether grouping pays for a batch (measured on MI355X, value + first
// partials, cubic): 3D tables of 17 MB and more gain from 2^20 points on
// (2^24 points: 1.4x at 128^3, 1.9x at 256^3); 2D tables up to 34 MB do
// not (their (k+1)-row stencils stay cache-friendly: 2.5x faster
// ungrouped at 1024^2 and 2^24 points), nor do 3D tables below 16 MB.
// 2D tables beyond the 256 MB Infinity Cache are grouped.
bool group_pays(int dim, int64_t table_bytes, int64_t npts) {
    if (npts < kGroupMinPoints) return false;
    if (dim == 3) return table_bytes >= (int64_t{16} << 20);
    return table_bytes >= (int64_t{256} << 20);
}

// Bricks of B^D stencil starts, B a power of two: the largest B whose
// (B + order)^D patch of table elements (elem_bytes each) stays within
// kPatchBytes (the LDS patch of the grouped
// evaluation kernel), coarsened until the brick key fits 16 bits (then
// the patch no longer fits and the grouped kernel reads HBM only).
BinParams bin_params(const GridData& g, int elem_bytes, int64_t* patch_volume) {
    BinParams bp{};
    bp.dim_count = g.dim;
    bp.order = g.order;
    int B = 1;
    auto patch = [&](int b) {
        int64_t v = 1;
        for (int d = 0; d < g.dim; ++d) v *= b + g.order;
        return v;
    };
    while (patch(2 * B) * elem_bytes <= kPatchBytes) B *= 2;
    for (;;) {
        int64_t nbin = 1;
        for (int d = 0; d < g.dim; ++d) nbin *= (g.setup[d].n + B - 1) / B;
        if (nbin <= 65536) break;
        B *= 2;
    }
    int shift = 0;
    while ((1 << shift) < B) ++shift;
    for (int d = 0; d < g.dim; ++d) {
        bp.dim[d] = dim_params(g, d);
        bp.bricks.brick_shift[d] = shift;
        bp.bricks.nbrick[d] = static_cast<int32_t>((g.setup[d].n + B - 1) / B);
    }
    *patch_volume = (patch(B) * elem_bytes <= kPatchBytes && B + g.order <= 64) ? patch(B) : 0;
    return bp;
}

// Bricks of the one-pass grouping: per dimension a power of two, grown
// round-robin from the last dimension while the (B_d + order) patch of
// elem_bytes elements fits kBrickPatchBytes (a brick no longer than its
// dimension, and a last-dimension patch row no longer than the kernel's
// workgroup).  False when that leaves more than kGroupMaxBins bricks.
bool item_bricks(const GridData& g, int elem_bytes, BrickGrid* out, int64_t* patch_volume, int32_t* nbins) {
    int shift[kGroupDim] = {0, 0, 0};
    auto volume = [&]() {
        int64_t v = 1;
        for (int d = 0; d < g.dim; ++d) v *= (int64_t{1} << shift[d]) + g.order;
        return v;
    };
    for (bool grew = true; grew;) {
        grew = false;
        for (int d = g.dim - 1; d >= 0; --d) {
            if ((int64_t{1} << (shift[d] + 1)) > g.setup[d].n) continue;
            if (d == g.dim - 1 && (int64_t{1} << (shift[d] + 1)) + g.order > kBrickThreads) continue;
            ++shift[d];
            if (volume() * elem_bytes <= kBrickPatchBytes)
                grew = true;
            else
                --shift[d];
        }
    }
    int64_t bins = 1;
    for (int d = 0; d < g.dim; ++d) {
        out->brick_shift[d] = shift[d];
        out->nbrick[d] = static_cast<int32_t>((g.setup[d].n + (int64_t{1} << shift[d]) - 1) >> shift[d]);
        bins *= out->nbrick[d];
    }
    if (bins > kGroupMaxBins || volume() * elem_bytes > kBrickPatchBytes) return false;
    *patch_volume = volume();
    *nbins = static_cast<int32_t>(bins);
    return true;
}

// Scratch bytes of the one-pass grouping (records, bin-major counts and
// their scan, scan tiles, brick items, item count) and of the two-pass
// grouping (two record arrays, histograms, keys).
int64_t one_pass_scratch(int64_t npts, int32_t nbins) {
    auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
    const int64_t nh = static_cast<int64_t>(nbins) * group_blocks(npts, nbins);
    return up(npts * static_cast<int64_t>(sizeof(BinRecord))) + 2 * up(nh * 4) + up(scan_tile_sums(nh) * 4) +
           up(group_max_items(npts, nbins) * static_cast<int64_t>(sizeof(BrickItem))) + 256;
}

int64_t two_pass_scratch(int64_t npts) {
    const int64_t nh = bin_hist_entries(npts), nt = (bin_scan_tiles(npts) + 3) / 4 * 4;
    return 2 * ((npts * static_cast<int64_t>(sizeof(BinRecord)) + 255) / 256 * 256) + (2 * nh + nt + (npts + 1) / 2) * 4 +
           256;
}

// The uniform fast path needs every dimension uniform (knots in closed
// form, basis polynomial classes uploaded) and 32-bit table offsets.
bool uniform_path(const GridData& g) {
    if (g.total >= (int64_t{1} << 32)) return false;
    for (int d = 0; d < g.dim; ++d)
        if (!g.setup[d].uniform || g.poly[d] == nullptr) return false;
    return true;
}

}  // namespace

// ---- the fixed scratch buffer of intp_eval_static ------------------------
//
// One file-scope device array; every GPU the library runs on holds its own
// copy (kStaticScratchBytes) until the process ends.  Host code never
// passes the array's name to a kernel: its device address on each GPU is
// resolved once per device with hipGetSymbolAddress (when a context is
// created on that device) and kept in g_static_scratch.
constexpr size_t kStaticScratchBytes = size_t{128} << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

std::mutex g_static_mu;
unsigned char* g_static_scratch[kMaxDevices] = {};

int resolve_static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(INTP_E_INVALID, "device index beyond the supported range");
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (g_static_scratch[device]) return INTP_OK;
    DeviceScope scope(device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress (static scratch)");
    g_static_scratch[device] = static_cast<unsigned char*>(ptr);
    return INTP_OK;
}

unsigned char* static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_static_mu);
    return device >= 0 && device < kMaxDevices ? g_static_scratch[device] : nullptr;
}

}  // namespace

extern "C" {

const char* intp_last_error(void) { return g_error.c_str(); }

const char* intp_status_string(int status) {
    switch (status) {
        case INTP_OK: return "ok";
        case INTP_E_INVALID: return "invalid argument";
        case INTP_E_DOMAIN: return "coordinate out of range";
        case INTP_E_NOMEM: return "out of device memory";
        case INTP_E_HIP: return "HIP runtime error";
        case INTP_E_UNSUPPORTED: return "unsupported dimension or order";
        case INTP_E_EVICTED: return "spline not resident";
        default: return "unknown status";
    }
}

int intp_ctx_create(int device, intp_ctx* out) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) return fail(INTP_E_INVALID, "no such device");
    if (!device_runtime(device)) return fail(INTP_E_HIP, "cannot create the device's streams");
    const int rc = resolve_static_scratch(device);
    if (rc != INTP_OK) return rc;
    *out = new intp_ctx_s{device};
    return INTP_OK;
}

int intp_ctx_destroy(intp_ctx ctx) {
    delete ctx;
    return INTP_OK;
}

int intp_template_create(intp_ctx ctx, const intp_grid_desc* desc, intp_template* out) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    std::shared_ptr<GridData> g;
    const int rc = make_grid(ctx, desc, nullptr, true, g);
    if (rc != INTP_OK) return rc;
    *out = new intp_template_s{std::move(g)};
    return INTP_OK;
}

int intp_template_destroy(intp_template t) {
    delete t;
    return INTP_OK;
}

}  // extern "C"

namespace {

constexpr int kMaxRefineSweeps = 3;

// Control points of `values` on a template's grid: FP64, or FP32 refined
// with `sweeps` FP64-residual corrections.
int interpolate_impl(intp_template t, const double* values, int values_mem, bool refined, int sweeps,
                     intp_spline* out, hipStream_t stream) {
    if (!t || !values || !out) return fail(INTP_E_INVALID, "null argument");
    if (values_mem != INTP_MEM_HOST && values_mem != INTP_MEM_DEVICE)
        return fail(INTP_E_INVALID, "values_mem must be INTP_MEM_HOST or INTP_MEM_DEVICE");
    if (refined && (sweeps < 0 || sweeps > kMaxRefineSweeps))
        return fail(INTP_E_INVALID, "refinement sweeps must be 0.." + std::to_string(kMaxRefineSweeps));
    const GridData& g = *t->grid;
    if (refined && g.total >= (int64_t{1} << 32))
        return fail(INTP_E_UNSUPPORTED, "refined builds need fewer than 2^32 control points");
    DeviceScope scope(g.device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    const size_t values_bytes = static_cast<size_t>(g.total) * sizeof(double);
    std::unique_ptr<intp_spline_s> s(new intp_spline_s());
    s->grid = t->grid;
    s->ctrl_bytes = refined ? 4 : 8;
    s->sweeps = refined ? sweeps : -1;
    const size_t ctrl_bytes = static_cast<size_t>(g.total) * s->ctrl_bytes;
    const size_t cells_at = (ctrl_bytes + 255) / 256 * 256;
    // gradient tables for 1 to 3 dimensions (cells of up to 4 floats)
    const bool with_cells = refined && g.dim <= kCellMaxDim;
    const size_t table_bytes = with_cells ? cells_at + static_cast<size_t>(g.total) * cell_bytes(g.dim) : ctrl_bytes;
    HIP_TRY(device_alloc(&s->ctrl, table_bytes, stream), "device allocation (control points)");
    if (with_cells) s->cells = static_cast<unsigned char*>(s->ctrl) + cells_at;
    const double* src = values;
    double* staged = nullptr;
    unsigned char* scratch = nullptr;
    int rc = INTP_OK;
    hipError_t e = hipSuccess;
    if (values_mem == INTP_MEM_HOST) {
        e = device_alloc(reinterpret_cast<void**>(&staged), values_bytes, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(staged, values, values_bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) rc = hip_fail(e, "staging values");
        src = staged;
    }
    if (rc == INTP_OK && refined && sweeps > 0) {
        const size_t bytes = ((static_cast<size_t>(g.total) * 4 + 255) / 256) * 256 +
                             2 * ((static_cast<size_t>(g.total) + 31) / 32) * 32 * sizeof(double);
        e = device_alloc(reinterpret_cast<void**>(&scratch), bytes, stream);
        if (e != hipSuccess) rc = hip_fail(e, "device allocation (refinement scratch)");
    }
    if (rc == INTP_OK)
        rc = refined ? refine_into(g, src, static_cast<float*>(s->ctrl), s->cells, sweeps, scratch, stream)
                     : solve_into(g, src, static_cast<double*>(s->ctrl), stream);
    // the spline's release waits for the solve; the grid's for its use here
    const hipError_t er = s->uses.record(stream);
    const hipError_t eg = t->grid->uses.record(stream);
    if (rc == INTP_OK && er != hipSuccess) rc = hip_fail(er, "hipEventRecord");
    if (rc == INTP_OK && eg != hipSuccess) rc = hip_fail(eg, "hipEventRecord");
    if (scratch) (void)hipFreeAsync(scratch, stream);
    if (staged) {
        (void)hipFreeAsync(staged, stream);
        const hipError_t es = hipStreamSynchronize(stream);  // the caller's host buffer must outlive the copy
        if (rc == INTP_OK && es != hipSuccess) rc = hip_fail(es, "hipStreamSynchronize");
    }
    if (rc != INTP_OK) return rc;
    *out = s.release();
    return INTP_OK;
}

}  // namespace

extern "C" {

int intp_interpolate(intp_template t, const double* values, int values_mem, intp_spline* out, void* stream) {
    return interpolate_impl(t, values, values_mem, false, 0, out, static_cast<hipStream_t>(stream));
}

int intp_interpolate_refined(intp_template t, const double* values, int values_mem, int32_t sweeps,
                             intp_spline* out, void* stream) {
    return interpolate_impl(t, values, values_mem, true, sweeps, out, static_cast<hipStream_t>(stream));
}

int intp_build_refined(intp_ctx ctx, const intp_grid_desc* desc, const double* values, int values_mem,
                       int32_t sweeps, intp_spline* out, void* stream) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    intp_template_s t;
    const int rc = build_grid(ctx, desc, static_cast<hipStream_t>(stream), t.grid);
    if (rc != INTP_OK) return rc;
    return intp_interpolate_refined(&t, values, values_mem, sweeps, out, stream);
}

int intp_build(intp_ctx ctx, const intp_grid_desc* desc, const double* values, int values_mem, intp_spline* out,
               void* stream) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    intp_template_s t;
    const int rc = build_grid(ctx, desc, static_cast<hipStream_t>(stream), t.grid);
    if (rc != INTP_OK) return rc;
    return intp_interpolate(&t, values, values_mem, out, stream);
}

int intp_spline_from_knots(intp_ctx ctx, int32_t dim, int32_t order, const int32_t* periodic, const int64_t* n,
                           const double* const* knots, const int64_t* nknots, const double* ctrl,
                           intp_spline* out) {
    if (!ctx || !n || !knots || !nknots || !ctrl || !out) return fail(INTP_E_INVALID, "null argument");
    int rc = check_dims(dim, order);
    if (rc != INTP_OK) return rc;
    auto g = std::make_shared<GridData>();
    g->device = ctx->device;
    g->dim = dim;
    g->order = order;
    for (int d = 0; d < dim; ++d) {
        if (!knots[d] || n[d] < 1) return fail(INTP_E_INVALID, "bad knot vector");
        const std::string err =
            setup_from_knots(order, periodic && periodic[d] != 0, n[d], knots[d], nknots[d], g->setup[d]);
        if (!err.empty()) return fail(INTP_E_INVALID, err + " (dimension " + std::to_string(d) + ")");
        g->total *= n[d];
    }
    DeviceScope scope(ctx->device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    DeviceRuntime* rt = device_runtime(ctx->device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    if ((rc = upload_grid(*g, rt->setup)) != INTP_OK) return rc;
    std::unique_ptr<intp_spline_s> s(new intp_spline_s());
    s->grid = std::move(g);
    const size_t bytes = static_cast<size_t>(s->grid->total) * sizeof(double);
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&s->ctrl), bytes, rt->setup), "device allocation (control points)");
    HIP_TRY(hipMemcpyAsync(s->ctrl, ctrl, bytes, hipMemcpyHostToDevice, rt->setup), "hipMemcpyAsync");
    HIP_TRY(s->uses.record(rt->setup), "hipEventRecord");
    HIP_TRY(hipStreamSynchronize(rt->setup), "hipStreamSynchronize");
    *out = s.release();
    return INTP_OK;
}

int intp_spline_retain(intp_spline s) {
    if (!s) return fail(INTP_E_INVALID, "null spline");
    s->refs.fetch_add(1);
    return INTP_OK;
}

int intp_spline_release(intp_spline s) {
    if (!s) return INTP_OK;
    if (s->refs.fetch_sub(1) == 1) delete s;
    return INTP_OK;
}

int intp_spline_get_info(intp_spline s, intp_spline_info* out) {
    if (!s || !out) return fail(INTP_E_INVALID, "null argument");
    const GridData& g = *s->grid;
    std::memset(out, 0, sizeof(*out));
    out->dim = g.dim;
    out->order = g.order;
    for (int d = 0; d < g.dim; ++d) {
        out->periodic[d] = g.setup[d].periodic;
        out->uniform[d] = g.setup[d].uniform;
        out->n[d] = g.setup[d].n;
        out->nknots[d] = static_cast<int64_t>(g.setup[d].knots.size());
        out->range_lo[d] = g.setup[d].lo;
        out->range_hi[d] = g.setup[d].hi;
    }
    out->ctrl_bytes = s->ctrl_bytes;
    out->refine_sweeps = s->sweeps;
    return INTP_OK;
}

int intp_spline_download_ctrl(intp_spline s, double* host, int64_t count) {
    if (!s || !host) return fail(INTP_E_INVALID, "null argument");
    if (count != s->grid->total) return fail(INTP_E_INVALID, "count must equal the number of control points");
    DeviceScope scope(s->grid->device);
    DeviceRuntime* rt = device_runtime(s->grid->device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    s->uses.order_before(rt->setup);
    if (s->ctrl_bytes == 4) {
        std::vector<float> tmp(static_cast<size_t>(count));
        HIP_TRY(hipMemcpyAsync(tmp.data(), s->ctrl, static_cast<size_t>(count) * sizeof(float),
                               hipMemcpyDeviceToHost, rt->setup),
                "hipMemcpyAsync");
        HIP_TRY(hipStreamSynchronize(rt->setup), "hipStreamSynchronize");
        for (int64_t i = 0; i < count; ++i) host[i] = static_cast<double>(tmp[static_cast<size_t>(i)]);
        return INTP_OK;
    }
    HIP_TRY(hipMemcpyAsync(host, s->ctrl, static_cast<size_t>(count) * sizeof(double), hipMemcpyDeviceToHost,
                           rt->setup),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(rt->setup), "hipStreamSynchronize");
    return INTP_OK;
}

int intp_spline_download_knots(intp_spline s, int32_t d, double* host, int64_t count) {
    if (!s || !host) return fail(INTP_E_INVALID, "null argument");
    if (d < 0 || d >= s->grid->dim) return fail(INTP_E_INVALID, "no such dimension");
    const auto& k = s->grid->setup[d].knots;
    if (count != static_cast<int64_t>(k.size())) return fail(INTP_E_INVALID, "count must equal the knot count");
    std::memcpy(host, k.data(), k.size() * sizeof(double));
    return INTP_OK;
}

}  // extern "C"

namespace {

// The gradient table's differences into s->mirror_diff (FP64).
int download_differences(intp_spline s) {
    const int D = s->grid->dim;
    const int64_t total = s->grid->total;
    const int per = cell_bytes(D) / static_cast<int>(sizeof(float));  // floats per cell
    DeviceScope scope(s->grid->device);
    DeviceRuntime* rt = device_runtime(s->grid->device);
    if (!rt) return fail(INTP_E_HIP, "cannot create the device's streams");
    s->uses.order_before(rt->setup);
    std::vector<float> cells(static_cast<size_t>(total) * per);
    HIP_TRY(hipMemcpyAsync(cells.data(), s->cells, cells.size() * sizeof(float), hipMemcpyDeviceToHost, rt->setup),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(rt->setup), "hipStreamSynchronize");
    std::vector<double> diff(static_cast<size_t>(total) * D);
    for (int d = 0; d < D; ++d)
        for (int64_t e = 0; e < total; ++e)
            diff[static_cast<size_t>(d * total + e)] = static_cast<double>(cells[static_cast<size_t>(e * per + 1 + d)]);
    s->mirror_diff = std::move(diff);
    return INTP_OK;
}

// The spline's control points on the host (FP64), downloaded once.
int ensure_mirror(intp_spline s) {
    if (s->mirror_ready.load(std::memory_order_acquire)) return INTP_OK;  // the per-call path
    std::call_once(s->mirror_once, [s]() {
        std::vector<double> host(static_cast<size_t>(s->grid->total));
        s->mirror_rc = intp_spline_download_ctrl(s, host.data(), s->grid->total);
        if (s->mirror_rc == INTP_OK && s->cells) s->mirror_rc = download_differences(s);
        if (s->mirror_rc == INTP_OK) {
            s->mirror = std::move(host);
            s->mirror_ready.store(true, std::memory_order_release);
        } else {
            s->mirror_error = g_error;
        }
    });
    if (s->mirror_rc != INTP_OK) return fail(s->mirror_rc, s->mirror_error);
    return INTP_OK;
}

HostSpline host_view(const intp_spline_s& s) {
    HostSpline h{};
    h.dim = s.grid->dim;
    h.order = s.grid->order;
    for (int d = 0; d < h.dim; ++d) h.dims[d] = &s.grid->setup[d];
    h.ctrl = s.mirror.data();
    h.diff = s.mirror_diff.empty() ? nullptr : s.mirror_diff.data();
    h.total = s.grid->total;
    return h;
}

int domain_error_at(const GridData& g, int64_t i, const double* x) {
    for (int d = 0; d < g.dim; ++d) {
        const DimSetup& ds = g.setup[d];
        if (!ds.periodic && !(x[d] >= ds.lo && x[d] <= ds.hi)) {
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "Given coordinate out of interpolation function range! (point %lld, dimension %d: %.17g "
                          "outside [%.17g, %.17g])",
                          static_cast<long long>(i), d, x[d], ds.lo, ds.hi);
            return fail(INTP_E_DOMAIN, buf);
        }
    }
    return fail(INTP_E_DOMAIN, "Given coordinate out of interpolation function range!");
}

// The first point of a batch with a clamped coordinate outside its range
// (host coordinates), or -1.
int64_t first_outside_host(const GridData& g, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt) {
    for (int64_t i = 0; i < npts; ++i)
        for (int d = 0; d < g.dim; ++d) {
            const DimSetup& ds = g.setup[d];
            const double x = coords[d * cs_dim + i * cs_pt];
            if (!ds.periodic && !(x >= ds.lo && x <= ds.hi)) return i;
        }
    return -1;
}

// Device scratch of one evaluation call (bounds flag, grouping records
// and tables): from the stream-ordered pool, or carved out of a fixed
// buffer that the caller owns for the call (intp_eval_static).
struct Scratch {
    hipStream_t stream = nullptr;
    unsigned char* base = nullptr;  // fixed buffer, or null: the pool
    size_t cap = 0, used = 0;

    bool fixed() const { return base != nullptr; }
    bool fits(size_t bytes) const { return !fixed() || (used + 255) / 256 * 256 + bytes <= cap; }
    hipError_t get(size_t bytes, void** out) {
        if (!fixed()) return device_alloc(out, bytes, stream);
        const size_t at = (used + 255) / 256 * 256;
        if (at + bytes > cap) return hipErrorOutOfMemory;
        *out = base + at;
        used = at + bytes;
        return hipSuccess;
    }
    void put(void* p) {
        if (!fixed() && p) (void)hipFreeAsync(p, stream);
    }
};

int eval_call(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt, int32_t nout,
              const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, uint32_t flags, hipStream_t stream,
              Scratch& sc);

}  // namespace

extern "C" {

int intp_eval(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt, int32_t nout,
              const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, uint32_t flags, void* stream_) {
    Scratch pool;
    pool.stream = static_cast<hipStream_t>(stream_);
    return eval_call(s, npts, coords, cs_dim, cs_pt, nout, orders, out, os_out, os_pt, flags, pool.stream, pool);
}

int intp_eval_static(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt, int32_t nout,
                     const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, uint32_t flags,
                     void* stream_) {
    if (!s) return fail(INTP_E_INVALID, "null spline");
    if (flags & (INTP_EVAL_HOST_MEM | INTP_EVAL_HOST_MIRROR | INTP_EVAL_TIMED))
        return fail(INTP_E_INVALID, "intp_eval_static takes device buffers only (no HOST_MEM, HOST_MIRROR or TIMED)");
    unsigned char* base = static_scratch(s->grid->device);
    if (!base) return fail(INTP_E_INVALID, "no context was created on the spline's device");
    Scratch fixed;
    fixed.stream = static_cast<hipStream_t>(stream_);
    fixed.base = base;
    fixed.cap = kStaticScratchBytes;
    return eval_call(s, npts, coords, cs_dim, cs_pt, nout, orders, out, os_out, os_pt, flags, fixed.stream, fixed);
}

int64_t intp_eval_static_capacity(void) { return static_cast<int64_t>(kStaticScratchBytes); }

}  // extern "C"

namespace {

int eval_call(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt, int32_t nout,
              const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, uint32_t flags, hipStream_t stream,
              Scratch& sc) {
    if (!s || !orders || nout < 1) return fail(INTP_E_INVALID, "null argument or no outputs");
    if (npts < 0) return fail(INTP_E_INVALID, "negative point count");
    if (npts == 0) return INTP_OK;
    if (!coords || !out) return fail(INTP_E_INVALID, "null coordinates or outputs");
    const GridData& g = *s->grid;
    const int D = g.dim;
    for (int i = 0; i < nout * D; ++i)
        if (orders[i] < 0) return fail(INTP_E_INVALID, "negative derivative order");
    if (cs_dim < 0 || cs_pt < 0 || os_out < 0 || os_pt < 0) return fail(INTP_E_INVALID, "negative stride");
    if ((flags & INTP_EVAL_GROUPED) && (flags & INTP_EVAL_UNGROUPED))
        return fail(INTP_E_INVALID, "INTP_EVAL_GROUPED and INTP_EVAL_UNGROUPED together");
    if (flags & INTP_EVAL_HOST_MIRROR) {
        // on this thread, from the host copy of the table
        if (flags & INTP_EVAL_OUT_F32) return fail(INTP_E_INVALID, "host evaluation writes FP64 results");
        int rc = ensure_mirror(s);
        if (rc != INTP_OK) return rc;
        const int64_t bad = host_eval_points(host_view(*s), npts, coords, cs_dim, cs_pt, nout, orders, out, os_out,
                                             os_pt, (flags & INTP_EVAL_BOUNDS_CHECK) != 0);
        if (bad >= 0) {
            double x[kMaxDim];
            for (int d = 0; d < D; ++d) x[d] = coords[d * cs_dim + bad * cs_pt];
            return domain_error_at(g, bad, x);
        }
        return INTP_OK;
    }
    DeviceScope scope(g.device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");

    EvalParams p{};
    for (int d = 0; d < D; ++d) p.dim[d] = dim_params(g, d);
    p.ctrl = s->ctrl;
    p.cs_dim = cs_dim;
    p.cs_pt = cs_pt;
    p.os_out = os_out;
    p.os_pt = os_pt;
    p.npts = npts;

    const bool host = (flags & INTP_EVAL_HOST_MEM) != 0;
    const bool out_f32 = (flags & INTP_EVAL_OUT_F32) != 0;
    const size_t out_elem = out_f32 ? sizeof(float) : sizeof(double);
    const size_t cext = static_cast<size_t>((D - 1) * cs_dim + (npts - 1) * cs_pt + 1);
    // Host outputs are staged in a dense device array [nout][npts] (or in
    // the caller's own layout when that is dense already) and copied back
    // once; nothing is copied to the device first.
    const bool dense_out = (os_pt == 1 && os_out == npts) || (os_out == 1 && os_pt == nout);
    const size_t oext = static_cast<size_t>(nout) * static_cast<size_t>(npts);
    double* dcoords = nullptr;
    void* dout = nullptr;
    int* dbad = nullptr;
    auto cleanup = [&]() {
        if (dcoords) (void)hipFreeAsync(dcoords, stream);
        if (dout) (void)hipFreeAsync(dout, stream);
        sc.put(dbad);
    };
    if (host) {
        hipError_t e = device_alloc(reinterpret_cast<void**>(&dcoords), cext * sizeof(double), stream);
        if (e == hipSuccess) e = device_alloc(&dout, oext * out_elem, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dcoords, coords, cext * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "staging host buffers");
        }
        p.coords = dcoords;
        if (!dense_out) {
            p.os_out = npts;
            p.os_pt = 1;
        }
    } else {
        p.coords = coords;
    }
    void* const out_dev = host ? dout : static_cast<void*>(out);
    if (out_f32)
        p.out32 = static_cast<float*>(out_dev);
    else
        p.out = static_cast<double*>(out_dev);
    if (flags & INTP_EVAL_BOUNDS_CHECK) {
        hipError_t e = sc.get(sizeof(int), reinterpret_cast<void**>(&dbad));
        // cleared by a kernel (graph-capturable; a fixed buffer keeps the
        // previous call's contents)
        if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(dbad), 0u, 1, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "bounds flag");
        }
        p.bad = dbad;
    }

    // value and first partials share one gradient launch; any other order
    // gets its own launch
    for (int i = 0; i <= kMaxDim; ++i) p.slot[i] = -1;
    std::vector<int> single;
    bool any_grad = false;
    for (int o = 0; o < nout; ++o) {
        const int32_t* r = orders + o * D;
        int sum = 0, which = -1;
        bool unit = true;
        for (int d = 0; d < D; ++d) {
            sum += r[d];
            if (r[d] == 1) which = d;
            if (r[d] > 1) unit = false;
        }
        int slot = -1;
        if (sum == 0) slot = 0;
        else if (sum == 1 && unit) slot = which + 1;
        if (slot >= 0 && p.slot[slot] < 0) {
            p.slot[slot] = o;
            any_grad = any_grad || slot > 0;
        } else {
            single.push_back(o);
        }
    }
    if (!any_grad && p.slot[0] >= 0) {
        // the value alone: the single-output kernel contracts one set of weights
        single.insert(single.begin(), p.slot[0]);
        p.slot[0] = -1;
    }
    {
        bool packed = !out_f32 && p.os_out == 1 && p.os_pt % 2 == 0 && reinterpret_cast<uintptr_t>(p.out) % 16 == 0;
        for (int i = 0; i <= D; ++i) packed = packed && p.slot[i] == i;
        p.out_packed = packed ? 1 : 0;
    }
    hipError_t e = hipSuccess;
    const bool fast = uniform_path(g);
    TimedCall timed{g.device, {nullptr, nullptr, nullptr}};
    const bool timing = (flags & INTP_EVAL_TIMED) != 0;
    if (timing) {
        for (int i = 0; i < 3 && e == hipSuccess; ++i) e = timing_event(g.device, &timed.ev[i]);
        if (e == hipSuccess) e = hipEventRecord(timed.ev[0], stream);
        if (e != hipSuccess) {
            for (hipEvent_t ev : timed.ev)
                if (ev) t_spare_events.emplace_back(g.device, ev);
            cleanup();
            return hip_fail(e, "timing events");
        }
    }
    BinRecord* rec_a = nullptr;
    BinRecord* rec_b = nullptr;
    uint32_t* bin_scratch = nullptr;
    auto free_bins = [&]() {
        sc.put(rec_a);
        sc.put(rec_b);
        sc.put(bin_scratch);
    };
    // refined builds: value and first partials come from the gradient table,
    // whose elements size the grouping's brick patches
    const bool cells = any_grad && s->cells != nullptr;
    const int elem_bytes = cells ? cell_bytes(D) : s->ctrl_bytes;
    const bool group = (flags & INTP_EVAL_GROUPED) ? npts > 0
                       : !(flags & INTP_EVAL_UNGROUPED) && group_pays(D, g.total * s->ctrl_bytes, npts);
    BrickGrid ib{};
    int64_t item_patch = 0;
    int32_t item_bins = 0;
    // Two radix passes by default (measured faster at 256^3); one pass on
    // request, and for a fixed scratch buffer, where it needs half the
    // memory.
    const bool groupable = fast && D >= 2 && D <= kGroupDim && group && npts <= 0xFFFFFFFFll;
    const bool bricks_ok = groupable && item_bricks(g, elem_bytes, &ib, &item_patch, &item_bins);
    const bool two_ok = groupable && bin_scan_tiles(npts) <= 4096;
    bool one_pass = bricks_ok && ((flags & INTP_EVAL_GROUP_ONE_PASS) || sc.fixed() || !two_ok);
    bool two_pass = !one_pass && two_ok;
    if (sc.fixed() && (one_pass || two_pass)) {
        // the grouping's scratch must fit the fixed buffer; otherwise the
        // points are evaluated in caller order (same results), unless the
        // caller forced grouping
        const size_t need = one_pass ? static_cast<size_t>(one_pass_scratch(npts, item_bins))
                                     : static_cast<size_t>(two_pass_scratch(npts));
        if (!sc.fits(need)) {
            if (flags & INTP_EVAL_GROUPED) {
                if (timing)
                    for (hipEvent_t ev : timed.ev) t_spare_events.emplace_back(g.device, ev);
                cleanup();
                return fail(INTP_E_INVALID, "grouping " + std::to_string(npts) + " points needs " +
                                                std::to_string(need) + " bytes of scratch; the static buffer holds " +
                                                std::to_string(sc.cap - sc.used));
            }
            one_pass = two_pass = false;
        }
    }
    if (one_pass) {
        GroupParams gp{};
        for (int d = 0; d < D; ++d) gp.dim[d] = p.dim[d];
        gp.bricks = ib;
        gp.dim_count = D;
        gp.order = g.order;
        gp.nbins = item_bins;
        gp.nblocks = static_cast<int32_t>(group_blocks(npts, item_bins));
        gp.coords = p.coords;
        gp.cs_dim = p.cs_dim;
        gp.cs_pt = p.cs_pt;
        gp.npts = npts;
        gp.max_items = group_max_items(npts, item_bins);
        const int64_t nh = static_cast<int64_t>(item_bins) * gp.nblocks;
        const int64_t nt = scan_tile_sums(nh);
        auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
        const int64_t b_rec = up(npts * static_cast<int64_t>(sizeof(BinRecord)));
        const int64_t b_hist = up(nh * 4), b_tiles = up(nt * 4);
        const int64_t b_items = up(gp.max_items * static_cast<int64_t>(sizeof(BrickItem)));
        e = sc.get(static_cast<size_t>(one_pass_scratch(npts, item_bins)), reinterpret_cast<void**>(&rec_a));
        if (e == hipSuccess) {
            unsigned char* base = reinterpret_cast<unsigned char*>(rec_a);
            gp.rec = rec_a;
            gp.hist = reinterpret_cast<uint32_t*>(base + b_rec);
            gp.offs = reinterpret_cast<uint32_t*>(base + b_rec + b_hist);
            gp.tile_sums = reinterpret_cast<uint32_t*>(base + b_rec + 2 * b_hist);
            gp.items = reinterpret_cast<BrickItem*>(base + b_rec + 2 * b_hist + b_tiles);
            gp.item_count = reinterpret_cast<uint32_t*>(base + b_rec + 2 * b_hist + b_tiles + b_items);
            e = launch_grouping(gp, stream);
        }
        if (e != hipSuccess) {
            free_bins();
            cleanup();
            if (timing)
                for (hipEvent_t ev : timed.ev) t_spare_events.emplace_back(g.device, ev);
            return hip_fail(e, "grouping query points");
        }
        p.coords = reinterpret_cast<const double*>(rec_a);
        p.cs_dim = 1;
        p.cs_pt = sizeof(BinRecord) / sizeof(double);
        p.bricks = ib;
        p.patch_volume = static_cast<int32_t>(item_patch);
        p.items = gp.items;
        p.item_count = gp.item_count;
        p.max_items = gp.max_items;
    } else if (two_pass) {
        int64_t patch_volume = 0;
        BinParams bp = bin_params(g, elem_bytes, &patch_volume);
        bp.coords = p.coords;
        bp.cs_dim = p.cs_dim;
        bp.cs_pt = p.cs_pt;
        bp.npts = npts;
        // histograms, their scan, scan tiles, then the 16-bit keys from a
        // 16-byte boundary (the pass-1 histogram reads them 8 at a time)
        const int64_t nh = bin_hist_entries(npts), nt = (bin_scan_tiles(npts) + 3) / 4 * 4;
        e = sc.get(static_cast<size_t>(npts) * sizeof(BinRecord), reinterpret_cast<void**>(&rec_a));
        if (e == hipSuccess) e = sc.get(static_cast<size_t>(npts) * sizeof(BinRecord), reinterpret_cast<void**>(&rec_b));
        if (e == hipSuccess)
            e = sc.get(static_cast<size_t>(2 * nh + nt + (npts + 1) / 2) * 4, reinterpret_cast<void**>(&bin_scratch));
        if (e == hipSuccess)
            e = launch_binning(bp, bin_scratch, bin_scratch + nh, bin_scratch + 2 * nh, bin_scratch + 2 * nh + nt,
                               rec_a, rec_b, stream);
        if (e != hipSuccess) {
            free_bins();
            cleanup();
            if (timing)
                for (hipEvent_t ev : timed.ev) t_spare_events.emplace_back(g.device, ev);
            return hip_fail(e, "grouping query points");
        }
        p.coords = reinterpret_cast<const double*>(rec_b);
        p.cs_dim = 1;
        p.cs_pt = sizeof(BinRecord) / sizeof(double);
        p.perm = reinterpret_cast<const uint32_t*>(rec_b) + offsetof(BinRecord, idx) / sizeof(uint32_t);
        p.perm_stride = sizeof(BinRecord) / sizeof(uint32_t);
        p.key = reinterpret_cast<const uint32_t*>(rec_b) + offsetof(BinRecord, key) / sizeof(uint32_t);
        p.bricks = bp.bricks;
        p.patch_volume = static_cast<int32_t>(patch_volume);
    }
    if (timing && e == hipSuccess) e = hipEventRecord(timed.ev[1], stream);
    const bool f32 = s->ctrl_bytes == 4;
    if (any_grad && e == hipSuccess) {
        if (cells) {
            EvalParams pc = p;
            pc.ctrl = s->cells;
            e = launch_eval_cells(D, g.order, fast, pc, stream);
        } else {
            e = launch_eval_grad(D, g.order, fast, f32, p, stream);
        }
    }
    for (size_t i = 0; e == hipSuccess && i < single.size(); ++i) {
        for (int d = 0; d < D; ++d) p.order[d] = orders[single[i] * D + d];
        p.out_index = single[i];
        e = launch_eval_single(D, g.order, fast, f32, p, stream);
    }
    if (timing) {
        if (e == hipSuccess) e = hipEventRecord(timed.ev[2], stream);
        if (e == hipSuccess) {
            t_timed.push_back(timed);
        } else {
            for (hipEvent_t ev : timed.ev) t_spare_events.emplace_back(g.device, ev);
        }
    }
    {
        // the spline's release waits for this call's kernels
        const hipError_t eu = s->uses.record(stream);
        if (e == hipSuccess) e = eu;
    }
    free_bins();
    if (e != hipSuccess) {
        cleanup();
        (void)hipStreamSynchronize(stream);
        if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "no kernel compiled for this dimension/order");
        return hip_fail(e, "eval launch");
    }
    int bad = 0;
    if (dbad) {
        e = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "bounds flag copy");
        }
    }
    std::vector<unsigned char> staged;
    if (host) {
        void* dst = out;
        if (!dense_out) {
            staged.resize(oext * out_elem);
            dst = staged.data();
        }
        e = hipMemcpyAsync(dst, dout, oext * out_elem, hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "output copy");
        }
    }
    cleanup();
    if (host || dbad) HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (host && !dense_out) {
        // [nout][npts] -> the caller's strides
        for (int o = 0; o < nout; ++o)
            for (int64_t i = 0; i < npts; ++i) {
                const size_t src = static_cast<size_t>(o) * static_cast<size_t>(npts) + static_cast<size_t>(i);
                if (out_f32)
                    reinterpret_cast<float*>(out)[o * os_out + i * os_pt] = reinterpret_cast<const float*>(staged.data())[src];
                else
                    out[o * os_out + i * os_pt] = reinterpret_cast<const double*>(staged.data())[src];
            }
    }
    if (bad) {
        // name the first offending point
        int64_t first = -1;
        if (host) {
            first = first_outside_host(g, npts, coords, cs_dim, cs_pt);
        } else {
            uint64_t* dfirst = nullptr;
            uint64_t hfirst = ~uint64_t{0};
            BoundsParams bp{};
            for (int d = 0; d < D; ++d) bp.dim[d] = p.dim[d];
            bp.ndim = D;
            bp.coords = coords;
            bp.cs_dim = cs_dim;
            bp.cs_pt = cs_pt;
            bp.npts = npts;
            sc.used = 0;  // the call's launches are done (synchronized above)
            e = sc.get(sizeof(uint64_t), reinterpret_cast<void**>(&dfirst));
            if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(dfirst), ~0u, 2, stream);
            bp.first = dfirst;
            if (e == hipSuccess) e = launch_first_outside(bp, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(&hfirst, dfirst, sizeof(uint64_t), hipMemcpyDeviceToHost, stream);
            sc.put(dfirst);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e == hipSuccess && hfirst != ~uint64_t{0}) first = static_cast<int64_t>(hfirst);
            if (first >= 0) {
                double x[kMaxDim] = {};
                for (int d = 0; d < D && e == hipSuccess; ++d)
                    e = hipMemcpy(&x[d], coords + d * cs_dim + first * cs_pt, sizeof(double), hipMemcpyDeviceToHost);
                if (e == hipSuccess) return domain_error_at(g, first, x);
            }
        }
        if (host && first >= 0) {
            double x[kMaxDim];
            for (int d = 0; d < D; ++d) x[d] = coords[d * cs_dim + first * cs_pt];
            return domain_error_at(g, first, x);
        }
        return fail(INTP_E_DOMAIN, "Given coordinate out of interpolation function range!");
    }
    return INTP_OK;
}

}  // namespace

struct intp_proxy_s {
    std::shared_ptr<GridData> grid;
    int64_t npts = 0;
    bool host = false;
    bool uniform = false;  // device weights from the uniform-grid path
    std::vector<int32_t> first_h;
    std::vector<double> weights_h;
    unsigned char* dev = nullptr;  // first (int32) then weights (double)
    int32_t* first_d = nullptr;
    double* weights_d = nullptr;
    UseTracker uses;
    ~intp_proxy_s() {
        if (dev) release_after(grid->device, dev, uses);
    }
};

namespace {

// Splines on equal grids: the same tables, or equal dimension, order,
// sizes, periodicity and knot vectors.
bool same_grid(const GridData& a, const GridData& b) {
    if (&a == &b) return true;
    if (a.dim != b.dim || a.order != b.order) return false;
    for (int d = 0; d < a.dim; ++d) {
        const DimSetup &x = a.setup[d], &y = b.setup[d];
        if (x.n != y.n || x.periodic != y.periodic || x.uniform != y.uniform || x.knots.size() != y.knots.size() ||
            std::memcmp(x.knots.data(), y.knots.data(), x.knots.size() * sizeof(double)) != 0)
            return false;
    }
    return true;
}

}  // namespace

extern "C" {

int intp_proxy_create(intp_template t, intp_spline s, int64_t npts, const double* coords, int64_t cs_dim,
                      int64_t cs_pt, uint32_t flags, intp_proxy* out, void* stream_) {
    if (!out || (!t == !s)) return fail(INTP_E_INVALID, "null output, or not exactly one of template and spline");
    if (npts < 0 || (npts > 0 && !coords)) return fail(INTP_E_INVALID, "null coordinates or negative point count");
    if (cs_dim < 0 || cs_pt < 0) return fail(INTP_E_INVALID, "negative stride");
    if (npts >= (int64_t{1} << 31)) return fail(INTP_E_UNSUPPORTED, "a proxy holds fewer than 2^31 points");
    const std::shared_ptr<GridData>& gp = t ? t->grid : s->grid;
    const GridData& g = *gp;
    const int D = g.dim, w = g.order + 1;
    std::unique_ptr<intp_proxy_s> p(new intp_proxy_s());
    p->grid = gp;
    p->npts = npts;
    p->host = (flags & INTP_EVAL_HOST_MEM) != 0;
    const bool bounds = (flags & INTP_EVAL_BOUNDS_CHECK) != 0;
    if (p->host) {
        p->first_h.resize(static_cast<size_t>(npts * D));
        p->weights_h.resize(static_cast<size_t>(npts * D * w));
        HostSpline hs{};
        hs.dim = D;
        hs.order = g.order;
        for (int d = 0; d < D; ++d) hs.dims[d] = &g.setup[d];
        const int64_t bad =
            host_proxy_weights(hs, npts, coords, cs_dim, cs_pt, bounds, p->first_h.data(), p->weights_h.data());
        if (bad >= 0) {
            double x[kMaxDim];
            for (int d = 0; d < D; ++d) x[d] = coords[d * cs_dim + bad * cs_pt];
            return domain_error_at(g, bad, x);
        }
        *out = p.release();
        return INTP_OK;
    }
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    DeviceScope scope(g.device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    const size_t first_bytes = (static_cast<size_t>(npts * D) * sizeof(int32_t) + 255) / 256 * 256;
    const size_t bytes = first_bytes + static_cast<size_t>(npts * D * w) * sizeof(double) + 256;
    HIP_TRY(device_alloc(reinterpret_cast<void**>(&p->dev), bytes, stream), "device allocation (proxy)");
    p->first_d = reinterpret_cast<int32_t*>(p->dev);
    p->weights_d = reinterpret_cast<double*>(p->dev + first_bytes);
    p->uniform = uniform_path(g);
    ProxyParams pp{};
    for (int d = 0; d < D; ++d) pp.dim[d] = dim_params(g, d);
    pp.npts = npts;
    pp.coords = coords;
    pp.cs_dim = cs_dim;
    pp.cs_pt = cs_pt;
    pp.first = p->first_d;
    pp.weights = p->weights_d;
    int* dbad = nullptr;
    hipError_t e = hipSuccess;
    if (bounds) {
        e = device_alloc(reinterpret_cast<void**>(&dbad), sizeof(int), stream);
        if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int), stream);
        pp.bad = dbad;
    }
    if (e == hipSuccess) e = launch_proxy_create(D, g.order, p->uniform, pp, stream);
    if (e == hipSuccess) e = p->uses.record(stream);
    if (e == hipSuccess) e = gp->uses.record(stream);
    int bad = 0;
    if (dbad) {
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, stream);
        (void)hipFreeAsync(dbad, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "no proxy kernel for this dimension/order");
    HIP_TRY(e, "proxy launch");
    if (bad) return fail(INTP_E_DOMAIN, "Given coordinate out of interpolation function range!");
    *out = p.release();
    return INTP_OK;
}

int intp_proxy_eval(intp_proxy p, intp_spline s, double* out, int64_t os_pt, uint32_t flags, void* stream_) {
    if (!p || !s) return fail(INTP_E_INVALID, "null proxy or spline");
    if (p->npts == 0) return INTP_OK;
    if (!out || os_pt < 0) return fail(INTP_E_INVALID, "null output or negative stride");
    if (!same_grid(*p->grid, *s->grid)) return fail(INTP_E_INVALID, "the spline is not on the proxy's grid");
    if (p->host) {
        if (flags & INTP_EVAL_OUT_F32) return fail(INTP_E_INVALID, "host proxies write FP64 results");
        int rc = ensure_mirror(s);
        if (rc != INTP_OK) return rc;
        host_proxy_apply(host_view(*s), p->npts, p->first_h.data(), p->weights_h.data(), out, os_pt);
        return INTP_OK;
    }
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    const GridData& g = *s->grid;
    DeviceScope scope(g.device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    ProxyParams pp{};
    for (int d = 0; d < g.dim; ++d) pp.dim[d] = dim_params(g, d);
    pp.npts = p->npts;
    pp.first = p->first_d;
    pp.weights = p->weights_d;
    pp.ctrl = s->ctrl;
    if (flags & INTP_EVAL_OUT_F32)
        pp.out32 = reinterpret_cast<float*>(out);
    else
        pp.out = out;
    pp.os_pt = os_pt;
    p->uses.order_before(stream);
    hipError_t e = launch_proxy_apply(g.dim, g.order, p->uniform, s->ctrl_bytes == 4, pp, stream);
    if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "no proxy kernel for this dimension/order");
    HIP_TRY(e, "proxy launch");
    HIP_TRY(p->uses.record(stream), "hipEventRecord");
    HIP_TRY(s->uses.record(stream), "hipEventRecord");
    return INTP_OK;
}

int intp_proxy_points(intp_proxy p, int64_t* npts) {
    if (!p || !npts) return fail(INTP_E_INVALID, "null argument");
    *npts = p->npts;
    return INTP_OK;
}

int intp_proxy_destroy(intp_proxy p) {
    delete p;
    return INTP_OK;
}

int intp_eval_timings(double* group_ms, double* kernel_ms, int64_t capacity, int64_t* count) {
    if (!count || capacity < 0 || (capacity > 0 && (!group_ms || !kernel_ms)))
        return fail(INTP_E_INVALID, "null argument or negative capacity");
    int rc = INTP_OK;
    int64_t n = 0;
    for (const TimedCall& t : t_timed) {
        DeviceScope scope(t.device);
        float g_ms = 0.f, k_ms = 0.f;
        hipError_t e = hipEventSynchronize(t.ev[2]);
        if (e == hipSuccess) e = hipEventElapsedTime(&g_ms, t.ev[0], t.ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&k_ms, t.ev[1], t.ev[2]);
        if (e != hipSuccess && rc == INTP_OK) rc = hip_fail(e, "timing events");
        if (n < capacity) {
            group_ms[n] = g_ms;
            kernel_ms[n] = k_ms;
        }
        ++n;
        for (hipEvent_t ev : t.ev) t_spare_events.emplace_back(t.device, ev);
    }
    t_timed.clear();
    *count = n;
    return rc;
}

}  // extern "C"
