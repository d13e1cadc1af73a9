// abi.hip -- the C ABI of libintp_amd.so (declared in include/intp_amd.h):
// handles, device memory, argument checks and kernel dispatch.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/intp_amd.h"
#include "kernels.hpp"
#include "setup.hpp"

using namespace intp_amd;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    const int code = e == hipErrorOutOfMemory ? INTP_E_NOMEM : INTP_E_HIP;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                    \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// Makes `device` current for the scope, restoring the caller's device.
class DeviceScope {
   public:
    explicit DeviceScope(int device) {
        ok_ = hipGetDevice(&prev_) == hipSuccess;
        if (ok_ && prev_ != device) ok_ = hipSetDevice(device) == hipSuccess;
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
    }
    bool ok() const { return ok_; }

   private:
    int prev_ = 0;
    bool ok_ = false;
};

struct DevArray {
    double* p = nullptr;
    size_t count = 0;
};

// Knots, factors and reciprocal tables of one grid on one device; shared
// by a template and every spline interpolated from it.
struct GridData {
    int device = 0;
    int dim = 0, order = 0;
    DimSetup setup[kMaxDim];
    DevArray knots[kMaxDim], recip[kMaxDim];
    DevArray lower[kMaxDim], upper[kMaxDim], diag[kMaxDim], bottom[kMaxDim], right[kMaxDim];
    DevArray poly[kMaxDim];
    bool has_factors = false;
    int64_t total = 1;

    ~GridData() {
        DeviceScope scope(device);
        for (int d = 0; d < kMaxDim; ++d)
            for (DevArray* a : {&knots[d], &recip[d], &lower[d], &upper[d], &diag[d], &bottom[d], &right[d],
                                &poly[d]})
                if (a->p) (void)hipFree(a->p);
    }
};

int upload(const std::vector<double>& v, DevArray& out) {
    out.count = v.size();
    if (v.empty()) return INTP_OK;
    HIP_TRY(hipMalloc(&out.p, v.size() * sizeof(double)), "hipMalloc");
    HIP_TRY(hipMemcpy(out.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    return INTP_OK;
}

int upload_grid(GridData& g) {
    for (int d = 0; d < g.dim; ++d) {
        DimSetup& s = g.setup[d];
        int rc;
        if ((rc = upload(s.knots, g.knots[d])) != INTP_OK) return rc;
        if ((rc = upload(s.recip, g.recip[d])) != INTP_OK) return rc;
        if ((rc = upload(s.poly, g.poly[d])) != INTP_OK) return rc;
        if (g.has_factors) {
            if ((rc = upload(s.lu.lower, g.lower[d])) != INTP_OK) return rc;
            if ((rc = upload(s.lu.upper, g.upper[d])) != INTP_OK) return rc;
            if ((rc = upload(s.lu.diag, g.diag[d])) != INTP_OK) return rc;
            if ((rc = upload(s.lu.bottom, g.bottom[d])) != INTP_OK) return rc;
            if ((rc = upload(s.lu.right, g.right[d])) != INTP_OK) return rc;
        }
    }
    return INTP_OK;
}

}  // namespace

struct intp_ctx_s {
    int device;
};

struct intp_template_s {
    std::shared_ptr<GridData> grid;
};

struct intp_spline_s {
    std::atomic<int> refs{1};
    std::shared_ptr<GridData> grid;
    double* ctrl = nullptr;
    ~intp_spline_s() {
        if (ctrl) {
            DeviceScope scope(grid->device);
            (void)hipFree(ctrl);
        }
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

int make_grid(intp_ctx ctx, const intp_grid_desc* desc, std::shared_ptr<GridData>& out) {
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
    if ((rc = upload_grid(*g)) != INTP_OK) return rc;
    out = std::move(g);
    return INTP_OK;
}

// Control-point solve of `values` (device) into `ctrl` (device):
// periodic shift copy, then one line-solve pass per axis, last axis first
// (InterpolationTemplate.hpp:448-580).
int solve_into(const GridData& g, const double* values, double* ctrl, hipStream_t stream) {
    ShiftParams sh{};
    sh.src = values;
    sh.dst = ctrl;
    sh.total = g.total;
    sh.dim = g.dim;
    bool any_shift = false;
    for (int d = 0; d < g.dim; ++d) {
        sh.n[d] = g.setup[d].n;
        sh.shift[d] = g.setup[d].periodic ? (g.order / 2) % g.setup[d].n : 0;
        any_shift = any_shift || sh.shift[d] != 0;
    }
    if (any_shift) {
        HIP_TRY(launch_shift_copy(sh, stream), "shift copy launch");
    } else if (values != ctrl) {
        HIP_TRY(hipMemcpyAsync(ctrl, values, static_cast<size_t>(g.total) * sizeof(double), hipMemcpyDeviceToDevice,
                               stream),
                "hipMemcpyAsync");
    }
    for (int a = g.dim - 1; a >= 0; --a) {
        const DimSetup& s = g.setup[a];
        SolveParams sp{};
        sp.x = ctrl;
        sp.n = s.n;
        int64_t inner = 1, outer = 1;
        for (int d = a + 1; d < g.dim; ++d) inner *= g.setup[d].n;
        for (int d = 0; d < a; ++d) outer *= g.setup[d].n;
        sp.inner = inner;
        sp.nlines = inner * outer;
        sp.lower = g.lower[a].p;
        sp.upper = g.upper[a].p;
        sp.diag = g.diag[a].p;
        sp.bottom = g.bottom[a].p;
        sp.right = g.right[a].p;
        hipError_t e = launch_solve(s.lu.p, s.lu.cyclic, sp, stream);
        if (e == hipErrorNotSupported) return fail(INTP_E_UNSUPPORTED, "band width without a compiled solver");
        HIP_TRY(e, "solve launch");
    }
    return INTP_OK;
}

DimParams dim_params(const GridData& g, int d) {
    const DimSetup& s = g.setup[d];
    DimParams p{};
    p.knots = g.knots[d].p;
    p.recip = g.recip[d].p;
    p.n = s.n;
    int64_t stride = 1;
    for (int e = d + 1; e < g.dim; ++e) stride *= g.setup[e].n;
    p.stride = stride;
    p.nt = static_cast<int32_t>(s.knots.size());
    p.periodic = s.periodic ? 1 : 0;
    p.uniform = s.uniform ? 1 : 0;
    p.lo = s.lo;
    p.hi = s.hi;
    p.period = s.hi - s.lo;
    p.inv_dx = s.uniform ? 1.0 / s.dx : 0.0;
    p.shift = s.periodic ? 1.0 : 0.5 * static_cast<double>(g.order + 1);
    p.xmin = s.xmin;
    p.xmax = s.xmax;
    p.dx = s.dx;
    p.half = s.half;
    p.poly = g.poly[d].p;
    p.nleft = s.nleft;
    p.nright = s.nright;
    p.nclass = s.nclass;
    return p;
}

// Batches at least this large are grouped by brick before evaluation.
constexpr int64_t kGroupMinPoints = int64_t{1} << 20;

// Brick of B^D control points per bin, B a power of two: the largest B
// whose (B + order)^D patch stays within 16 KB (well inside a CU's L1),
// coarsened until the brick key fits 16 bits.
BinParams bin_params(const GridData& g) {
    BinParams bp{};
    bp.dim_count = g.dim;
    int B = 1;
    auto patch = [&](int b) {
        double v = 8.0;
        for (int d = 0; d < g.dim; ++d) v *= static_cast<double>(b + g.order);
        return v;
    };
    while (patch(2 * B) <= 16384.0) B *= 2;
    for (;;) {
        int64_t nbin = 1;
        for (int d = 0; d < g.dim; ++d) nbin *= (g.setup[d].n + B - 1) / B;
        if (nbin <= 65536) break;
        B *= 2;
    }
    for (int d = 0; d < g.dim; ++d) {
        const DimSetup& s = g.setup[d];
        BinDim& b = bp.dim[d];
        b.lo = s.lo;
        b.period = s.hi - s.lo;
        b.inv_period = 1.0 / b.period;
        b.inv_dx = 1.0 / s.dx;
        b.periodic = s.periodic ? 1 : 0;
        b.ncell = static_cast<int32_t>(s.n);
        int shift = 0;
        while ((1 << shift) < B) ++shift;
        b.brick_shift = shift;
        b.nbrick = static_cast<int32_t>((s.n + B - 1) / B);
    }
    return bp;
}

// The uniform fast path needs every dimension uniform (knots in closed
// form, basis polynomial classes uploaded) and 32-bit table offsets.
bool uniform_path(const GridData& g) {
    if (g.total >= (int64_t{1} << 32)) return false;
    for (int d = 0; d < g.dim; ++d)
        if (!g.setup[d].uniform || g.poly[d].p == nullptr) return false;
    return true;
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
        default: return "unknown status";
    }
}

int intp_ctx_create(int device, intp_ctx* out) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) return fail(INTP_E_INVALID, "no such device");
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
    const int rc = make_grid(ctx, desc, g);
    if (rc != INTP_OK) return rc;
    *out = new intp_template_s{std::move(g)};
    return INTP_OK;
}

int intp_template_destroy(intp_template t) {
    delete t;
    return INTP_OK;
}

int intp_interpolate(intp_template t, const double* values, int values_mem, intp_spline* out, void* stream_) {
    if (!t || !values || !out) return fail(INTP_E_INVALID, "null argument");
    const GridData& g = *t->grid;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    DeviceScope scope(g.device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    const size_t bytes = static_cast<size_t>(g.total) * sizeof(double);
    std::unique_ptr<intp_spline_s> s(new intp_spline_s());
    s->grid = t->grid;
    HIP_TRY(hipMalloc(&s->ctrl, bytes), "hipMalloc (control points)");
    const double* src = values;
    double* staged = nullptr;
    if (values_mem == INTP_MEM_HOST) {
        HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&staged), bytes, stream), "hipMallocAsync");
        HIP_TRY(hipMemcpyAsync(staged, values, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
        src = staged;
    } else if (values_mem != INTP_MEM_DEVICE) {
        return fail(INTP_E_INVALID, "values_mem must be INTP_MEM_HOST or INTP_MEM_DEVICE");
    }
    int rc = solve_into(g, src, s->ctrl, stream);
    if (staged) {
        if (rc == INTP_OK) {
            HIP_TRY(hipFreeAsync(staged, stream), "hipFreeAsync");
            HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        } else {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(staged);
        }
    }
    if (rc != INTP_OK) return rc;
    *out = s.release();
    return INTP_OK;
}

int intp_build(intp_ctx ctx, const intp_grid_desc* desc, const double* values, int values_mem, intp_spline* out,
               void* stream) {
    if (!out) return fail(INTP_E_INVALID, "null output handle");
    intp_template_s t;
    const int rc = make_grid(ctx, desc, t.grid);
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
    if ((rc = upload_grid(*g)) != INTP_OK) return rc;
    std::unique_ptr<intp_spline_s> s(new intp_spline_s());
    s->grid = std::move(g);
    const size_t bytes = static_cast<size_t>(s->grid->total) * sizeof(double);
    HIP_TRY(hipMalloc(&s->ctrl, bytes), "hipMalloc (control points)");
    HIP_TRY(hipMemcpy(s->ctrl, ctrl, bytes, hipMemcpyHostToDevice), "hipMemcpy");
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
    return INTP_OK;
}

int intp_spline_download_ctrl(intp_spline s, double* host, int64_t count) {
    if (!s || !host) return fail(INTP_E_INVALID, "null argument");
    if (count != s->grid->total) return fail(INTP_E_INVALID, "count must equal the number of control points");
    DeviceScope scope(s->grid->device);
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_TRY(hipMemcpy(host, s->ctrl, static_cast<size_t>(count) * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
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

int intp_eval(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt, int32_t nout,
              const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, uint32_t flags, void* stream_) {
    if (!s || !orders || nout < 1) return fail(INTP_E_INVALID, "null argument or no outputs");
    if (npts < 0) return fail(INTP_E_INVALID, "negative point count");
    if (npts == 0) return INTP_OK;
    if (!coords || !out) return fail(INTP_E_INVALID, "null coordinates or outputs");
    const GridData& g = *s->grid;
    const int D = g.dim;
    for (int i = 0; i < nout * D; ++i)
        if (orders[i] < 0) return fail(INTP_E_INVALID, "negative derivative order");
    if (cs_dim < 0 || cs_pt < 0 || os_out < 0 || os_pt < 0) return fail(INTP_E_INVALID, "negative stride");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
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
    const size_t cext = static_cast<size_t>((D - 1) * cs_dim + (npts - 1) * cs_pt + 1);
    const size_t oext = static_cast<size_t>((nout - 1) * os_out + (npts - 1) * os_pt + 1);
    double* dcoords = nullptr;
    double* dout = nullptr;
    int* dbad = nullptr;
    auto cleanup = [&]() {
        if (dcoords) (void)hipFreeAsync(dcoords, stream);
        if (dout) (void)hipFreeAsync(dout, stream);
        if (dbad) (void)hipFreeAsync(dbad, stream);
    };
    if (host) {
        hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&dcoords), cext * sizeof(double), stream);
        if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&dout), oext * sizeof(double), stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dcoords, coords, cext * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dout, out, oext * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "staging host buffers");
        }
        p.coords = dcoords;
        p.out = dout;
    } else {
        p.coords = coords;
        p.out = out;
    }
    if (flags & INTP_EVAL_BOUNDS_CHECK) {
        hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&dbad), sizeof(int), stream);
        if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int), stream);
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
            any_grad = true;
        } else {
            single.push_back(o);
        }
    }
    hipError_t e = hipSuccess;
    const bool fast = uniform_path(g);
    BinRecord* rec_a = nullptr;
    BinRecord* rec_b = nullptr;
    uint32_t* bin_scratch = nullptr;
    auto free_bins = [&]() {
        if (rec_a) (void)hipFreeAsync(rec_a, stream);
        if (rec_b) (void)hipFreeAsync(rec_b, stream);
        if (bin_scratch) (void)hipFreeAsync(bin_scratch, stream);
    };
    if (fast && D >= 2 && npts >= kGroupMinPoints && npts <= 0xFFFFFFFFll && !(flags & INTP_EVAL_UNGROUPED) &&
        bin_scan_tiles(npts) <= 4096) {
        BinParams bp = bin_params(g);
        bp.coords = p.coords;
        bp.cs_dim = p.cs_dim;
        bp.cs_pt = p.cs_pt;
        bp.npts = npts;
        const int64_t nh = bin_hist_entries(npts), nt = bin_scan_tiles(npts);
        e = hipMallocAsync(reinterpret_cast<void**>(&rec_a), static_cast<size_t>(npts) * sizeof(BinRecord), stream);
        if (e == hipSuccess)
            e = hipMallocAsync(reinterpret_cast<void**>(&rec_b), static_cast<size_t>(npts) * sizeof(BinRecord),
                               stream);
        if (e == hipSuccess)
            e = hipMallocAsync(reinterpret_cast<void**>(&bin_scratch), static_cast<size_t>(2 * nh + nt + npts) * 4,
                               stream);
        if (e == hipSuccess)
            e = launch_binning(bp, bin_scratch, bin_scratch + nh, bin_scratch + 2 * nh, bin_scratch + 2 * nh + nt,
                               rec_a, rec_b, stream);
        if (e != hipSuccess) {
            free_bins();
            cleanup();
            return hip_fail(e, "grouping query points");
        }
        p.coords = reinterpret_cast<const double*>(rec_b);
        p.cs_dim = 1;
        p.cs_pt = sizeof(BinRecord) / sizeof(double);
        p.perm = reinterpret_cast<const uint32_t*>(rec_b) + offsetof(BinRecord, idx) / sizeof(uint32_t);
        p.perm_stride = sizeof(BinRecord) / sizeof(uint32_t);
    }
    if (any_grad) e = launch_eval_grad(D, g.order, fast, p, stream);
    for (size_t i = 0; e == hipSuccess && i < single.size(); ++i) {
        for (int d = 0; d < D; ++d) p.order[d] = orders[single[i] * D + d];
        p.out_index = single[i];
        e = launch_eval_single(D, g.order, fast, p, stream);
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
    if (host) {
        e = hipMemcpyAsync(out, dout, oext * sizeof(double), hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "output copy");
        }
    }
    cleanup();
    if (host || dbad) HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (bad) return fail(INTP_E_DOMAIN, "Given coordinate out of interpolation function range!");
    return INTP_OK;
}

}  // extern "C"
