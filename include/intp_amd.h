/*
 * intp_amd.h -- C ABI of the MI355X (gfx950) B-spline interpolation library
 * libintp_amd.so.  Plain pointers and sizes; no C++ or torch types.
 *
 * The reference (12ff54e/BSplineInterpolation) is a header-only C++ template
 * library with no C ABI of its own; each entry point below names the
 * reference member it replaces.  The C++ drop-in header
 * include/intp/Interpolation.hpp rebuilds the reference's class surface on
 * top of these calls.
 *
 * Conventions (the reference's test configuration, test/CMakeLists.txt:46-47):
 *  - periodic dimensions store n values with spacing (xmax - xmin)/n
 *    (INTP_PERIODIC_NO_DUMMY_POINT);
 *  - value meshes are C-order (last index fastest), Mesh.hpp:87-95;
 *  - every call returns an intp_status; intp_last_error() describes the
 *    last failure on the calling thread.
 *
 * Thread safety: splines and templates are immutable after creation, so
 * any number of threads may evaluate one spline concurrently, on any
 * streams (the reference's evaluation members are const and reentrant,
 * Interpolation.hpp:132-244).  Creation and release are thread-safe.
 */
#ifndef INTP_AMD_H
#define INTP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INTP_MAX_DIM 6   /* dimensions with compiled kernels (batches: 3) */
#define INTP_MAX_ORDER 7 /* spline orders with compiled kernels         */

typedef enum {
    INTP_OK = 0,
    INTP_E_INVALID = 1,     /* bad argument; reference: std::runtime_error
                               from INTP_ASSERT (util.hpp:247-258)        */
    INTP_E_DOMAIN = 2,      /* coordinate out of range; reference:
                               std::domain_error (Interpolation.hpp:482-491) */
    INTP_E_NOMEM = 3,       /* device allocation failed; std::bad_alloc  */
    INTP_E_HIP = 4,         /* HIP runtime error                         */
    INTP_E_UNSUPPORTED = 5, /* dimension/order without compiled kernels  */
    INTP_E_EVICTED = 6      /* serving: the spline is no longer resident
                               (evicted, released or its build failed)    */
} intp_status;

/* where a caller's buffer lives */
#define INTP_MEM_DEVICE 0
#define INTP_MEM_HOST 1

/* eval flags */
#define INTP_EVAL_BOUNDS_CHECK 1u /* at() / derivative_at() semantics     */
#define INTP_EVAL_HOST_MEM 2u     /* coords and outputs are host memory   */
#define INTP_EVAL_UNGROUPED 4u    /* evaluate in the caller's point order:
                                     no grouping of large batches by brick */
#define INTP_EVAL_GROUPED 32u     /* group by brick whenever the grid
                                     allows it (uniform 2D/3D grids),
                                     whatever the batch and table sizes    */
#define INTP_EVAL_TIMED 8u        /* record HIP events around this call's
                                     grouping and its evaluation kernels,
                                     read back with intp_eval_timings()   */
#define INTP_EVAL_OUT_F32 16u     /* `out` points to float storage: results
                                     are stored as FP32 (refined splines) */
#define INTP_EVAL_GROUP_ONE_PASS 128u /* group in one counting pass over
                                     large bricks (64 KB patches, brick items)
                                     where the grid allows it, instead of two
                                     radix passes over small bricks: half the
                                     scratch memory (32 bytes per point), a
                                     few percent slower at 256^3 (random
                                     record writes)                        */
#define INTP_EVAL_HOST_MIRROR 64u /* evaluate on the calling host thread from
                                     a host copy of the control points
                                     (downloaded on the first such call and
                                     kept with the spline); coords and out
                                     are host memory, the call is
                                     synchronous and launches nothing.  The
                                     scalar members of the C++ drop-in use
                                     it: one point per call, where a device
                                     round trip costs tens of microseconds */

typedef struct intp_ctx_s* intp_ctx;
typedef struct intp_template_s* intp_template;
typedef struct intp_spline_s* intp_spline;
typedef struct intp_proxy_s* intp_proxy;

/* Grid description: one InterpolationFunctionTemplate's arguments
 * (InterpolationTemplate.hpp:60-79; Interpolation.hpp:92-107).
 * coords[d] == NULL: uniform dimension on [xmin[d], xmax[d]]
 *   (the std::pair<T,T>(min, max) range, Interpolation.hpp:322-362);
 * coords[d] != NULL: host array of n[d] + periodic[d] node coordinates
 *   (the iterator-pair range, Interpolation.hpp:365-464).               */
typedef struct {
    int32_t dim;
    int32_t order;
    int32_t periodic[INTP_MAX_DIM];
    int64_t n[INTP_MAX_DIM];
    double xmin[INTP_MAX_DIM];
    double xmax[INTP_MAX_DIM];
    const double* coords[INTP_MAX_DIM];
} intp_grid_desc;

/* Properties of a spline: InterpolationFunction::periodicity / uniform /
 * range / get_order (Interpolation.hpp:248-267).                         */
typedef struct {
    int32_t dim;
    int32_t order;
    int32_t periodic[INTP_MAX_DIM];
    int32_t uniform[INTP_MAX_DIM];
    int64_t n[INTP_MAX_DIM];  /* control points per dimension */
    int64_t nknots[INTP_MAX_DIM];
    double range_lo[INTP_MAX_DIM];
    double range_hi[INTP_MAX_DIM];
    int32_t ctrl_bytes;    /* 8: FP64 control points; 4: refined FP32      */
    int32_t refine_sweeps; /* FP64-residual sweeps of a refined build, else -1 */
} intp_spline_info;

/* ---- errors -------------------------------------------------------- */
const char* intp_last_error(void);
const char* intp_status_string(int status);

/* ---- context ------------------------------------------------------- */
/* Binds a device (hipSetDevice) for every later call made with objects of
 * this context.  No reference counterpart (the reference has no device). */
int intp_ctx_create(int device, intp_ctx* out);
int intp_ctx_destroy(intp_ctx ctx);

/* ---- template: knots + factored collocation matrices ---------------- */
/* Replaces InterpolationFunctionTemplate<T,D,O,U>(periodicity,
 * MeshDimension, x_ranges...) (InterpolationTemplate.hpp:60-79): builds
 * knot vectors (Interpolation.hpp:322-480), assembles and LU-factors the
 * per-dimension collocation matrices (InterpolationTemplate.hpp:254-446,
 * BandLU.hpp:103-213) and uploads them.                                  */
int intp_template_create(intp_ctx ctx, const intp_grid_desc* desc, intp_template* out);
int intp_template_destroy(intp_template t);

/* Replaces InterpolationFunctionTemplate::interpolate(mesh)
 * (InterpolationTemplate.hpp:118-143, 448-580): solves for the control
 * points of `values` (C-order, n[0] x ... x n[dim-1] doubles) on the GPU
 * and returns a new spline holding them in HBM.  values_mem is
 * INTP_MEM_HOST or INTP_MEM_DEVICE.  Asynchronous on `stream`
 * (hipStream_t, NULL = default) for device input.                       */
int intp_interpolate(intp_template t, const double* values, int values_mem, intp_spline* out,
                     void* stream);

/* Replaces InterpolationFunction<T,D,O,U>(periodicity, f_mesh,
 * x_ranges...) (Interpolation.hpp:92-107): template + interpolate.      */
int intp_build(intp_ctx ctx, const intp_grid_desc* desc, const double* values, int values_mem,
               intp_spline* out, void* stream);

/* Refined-FP32 form of intp_interpolate / intp_build: control points
 * stored in FP32, solved with the FP32-rounded LU factors (FP64 factors
 * for periodic dimensions of order >= 5), then `sweeps` (0..3) rounds of
 * iterative refinement: the residual r = S f - (M_0 (x) ... (x) M_{D-1}) c
 * is formed in FP64 from the unfactored collocation matrices and the FP64
 * values, solved in FP32 and added to c.  Splines built this way are
 * evaluated with FP32 weights and tables.  A refined spline also keeps a
 * gradient table (8 bytes per control point in 1D, 16 in 2D and 3D,
 * beside the 4 of the FP32 control points): each FP32 control point with
 * its backward differences along each dimension,
 * formed in FP64 from c plus the last round's correction; value and
 * first partials are evaluated from it, so that the partials carry FP32
 * rounding of the differences rather than cancellation of FP32 control
 * points (about 1e-7 relative instead of 5e-6 with one sweep or more;
 * with 0 sweeps the differences come from the FP32 solve alone).  Host
 * evaluation (INTP_EVAL_HOST_MIRROR) reads the same differences for first
 * partials.  Splines of 4 to 6 dimensions keep no gradient table.  No
 * reference counterpart: the reference solves in the value type T
 * (InterpolationTemplate.hpp:448-580); this is its mixed-precision
 * variant.                                                              */
int intp_interpolate_refined(intp_template t, const double* values, int values_mem, int32_t sweeps,
                             intp_spline* out, void* stream);
int intp_build_refined(intp_ctx ctx, const intp_grid_desc* desc, const double* values, int values_mem,
                       int32_t sweeps, intp_spline* out, void* stream);

/* Replaces BSpline<T,D,O,U>(periodicity, ctrl_pts, knot_pairs...)
 * (BSpline.hpp:188-210): a spline from explicit knot vectors (host arrays,
 * nknots[d] each) and control points (host, C-order, n[d] each with
 * nknots[d] = n[d] + order + 1 (+order if periodic)).                    */
int intp_spline_from_knots(intp_ctx ctx, int32_t dim, int32_t order, const int32_t* periodic,
                           const int64_t* n, const double* const* knots, const int64_t* nknots,
                           const double* ctrl, intp_spline* out);

/* Reference counting (value semantics of the reference: a copy of a built
 * spline shares its immutable device table).                            */
int intp_spline_retain(intp_spline s);
int intp_spline_release(intp_spline s);
int intp_spline_get_info(intp_spline s, intp_spline_info* out);
/* Copies the control points (C-order) to host memory; synchronous.       */
int intp_spline_download_ctrl(intp_spline s, double* host, int64_t count);
/* Copies knot vector `d` to host memory (nknots[d] doubles).             */
int intp_spline_download_knots(intp_spline s, int32_t d, double* host, int64_t count);

/* ---- evaluation ---------------------------------------------------- */
/* Replaces operator() / derivative() and, with INTP_EVAL_BOUNDS_CHECK,
 * at() / derivative_at() (Interpolation.hpp:132-244; BSpline.hpp:305-550),
 * batched over points:
 *   coordinate d of point i : coords[d * cs_dim + i * cs_pt]
 *   output o of point i     : out[o * os_out + i * os_pt]
 *   derivative orders       : orders[o * dim + d], o < nout
 * (SoA: cs_dim = npts, cs_pt = 1; AoS: cs_dim = 1, cs_pt = dim.)
 * All outputs of one call share one stencil gather.  Out-of-range
 * coordinates of clamped dimensions extrapolate the boundary polynomial,
 * periodic ones wrap; with INTP_EVAL_BOUNDS_CHECK the call returns
 * INTP_E_DOMAIN (and synchronizes `stream`) if any clamped coordinate is
 * outside range(d).  Orders above the spline order give 0.
 * Large batches on uniform 2D/3D grids are first grouped by brick of the
 * control-point table so that each workgroup gathers from a patch of the
 * table staged in LDS: two radix passes over 32-byte records (scratch
 * device memory from the stream-ordered allocator, about 68 bytes per
 * point), or with INTP_EVAL_GROUP_ONE_PASS one counting pass over larger
 * bricks (about 32 bytes per point).
 * Results are written to the caller's positions and do not depend on the
 * grouping.                                                              */
int intp_eval(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
              int32_t nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt,
              uint32_t flags, void* stream);

/* intp_eval without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, kernels and results as intp_eval, but every scratch buffer
 * (the bounds flag, the grouping's records and tables) is carved from one
 * fixed device buffer per GPU of intp_eval_static_capacity() bytes
 * (128 MiB), held from library load until the process ends; the call
 * makes no hipMalloc/hipFree or other allocation.  Device buffers only:
 * INTP_EVAL_HOST_MEM, INTP_EVAL_HOST_MIRROR and INTP_EVAL_TIMED are
 * refused.  A batch whose grouping would not fit the buffer is evaluated
 * in caller order (same results); with INTP_EVAL_GROUPED it is refused
 * (INTP_E_INVALID).  The buffer is shared by every call on that GPU and
 * keeps the last call's contents: calls that overlap on one GPU (other
 * threads, other streams) race and must be serialized by the caller.  A
 * context must have been created on the spline's device.                */
int intp_eval_static(intp_spline s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                     int32_t nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt,
                     uint32_t flags, void* stream);
int64_t intp_eval_static_capacity(void);

/* With INTP_EVAL_BOUNDS_CHECK, INTP_E_DOMAIN names the first offending
 * point (index, dimension, coordinate and range) in intp_last_error();
 * the outputs of a batch that fails the check are unspecified.           */

/* ---- evaluation proxies ------------------------------------------------ */
/* Replaces eval_proxy / pre_calc_coef (Interpolation.hpp:493-506,
 * InterpolationTemplate.hpp:145-176, BSpline.hpp:244-297), for many points
 * at once: the knot spans and value basis of `npts` fixed points on the
 * grid of template `t` or of spline `s` (exactly one of them) are computed
 * once; intp_proxy_eval then gives each point's value for any spline on an
 * equal grid with one gather and dot product per point, no span search and
 * no basis evaluation.
 *   flags & INTP_EVAL_HOST_MEM: a host proxy (coords host memory; applied
 *     on the calling thread to the spline's host copy of its control
 *     points, out host memory): the reference's per-point proxy;
 *   otherwise a device proxy (coords device memory, asynchronous on
 *     `stream`; applied by a kernel, out device memory, FP32 with
 *     INTP_EVAL_OUT_F32): its values equal intp_eval's bit for bit.
 *   flags & INTP_EVAL_BOUNDS_CHECK: at() semantics at creation.
 * Output of point i: out[i * os_pt].                                      */
int intp_proxy_create(intp_template t, intp_spline s, int64_t npts, const double* coords, int64_t cs_dim,
                      int64_t cs_pt, uint32_t flags, intp_proxy* out, void* stream);
int intp_proxy_eval(intp_proxy p, intp_spline s, double* out, int64_t os_pt, uint32_t flags, void* stream);
int intp_proxy_points(intp_proxy p, int64_t* npts);
int intp_proxy_destroy(intp_proxy p);

/* ---- batches: many independent splines built and evaluated together -- */
/* No single reference counterpart: a batch is N independent
 * InterpolationFunction<double, D_s, O_s> objects (Interpolation.hpp:92-107)
 * of mixed dimension (1 to 3), order, size and periodicity, each built and evaluated
 * with the single-spline arithmetic (its control points equal
 * intp_build's), but by a fixed set of persistent launches over all of
 * them.  Layouts:
 *   values / control points: the splines' C-order meshes concatenated in
 *     spline order (intp_batch_info gives the offsets);
 *   points: coords[i*3 + d] (x, y, z; unused dimensions ignored), spline s
 *     owning points point_offsets[s] .. point_offsets[s+1]-1 (host array of
 *     count + 1); out[i] its value.                                        */
typedef struct intp_batch_s* intp_batch;
#define INTP_BATCH_GRAPH 1u    /* capture build + eval into one hipGraph and
                                  replay it while the buffers, the stream and
                                  the point layout stay the same (stream must
                                  not be the null stream)                  */
#define INTP_BATCH_AUTOTUNE 2u /* on the first run: time candidate tile
                                  shapes (solve lines per tile, persistent
                                  workgroups per CU, points per eval
                                  workgroup) on this batch, keep the fastest */
#define INTP_BATCH_SPLIT_LINES 4u /* solve 1D splines of >= 256 points as
                                  up to 63 segments in parallel, chained by
                                  precomputed segment responses: the same
                                  linear system, but rounded differently
                                  (within 1e-12 relative of the sequential
                                  solve instead of bit-identical)          */
int intp_batch_create(intp_ctx ctx, int32_t count, const intp_grid_desc* descs, intp_batch* out);
int intp_batch_destroy(intp_batch b);
/* Total values (= control points) and the per-spline offsets (count + 1). */
int intp_batch_info(intp_batch b, int64_t* values_total, int64_t* value_offsets);
int intp_batch_build(intp_batch b, const double* values, void* stream);
int intp_batch_eval(intp_batch b, const double* coords, const int64_t* point_offsets, double* out, void* stream);
/* Build then evaluate (values, coords, out: device). */
int intp_batch_run(intp_batch b, const double* values, const double* coords, const int64_t* point_offsets,
                   double* out, uint32_t flags, void* stream);
/* The tile shapes in use and the autotuner's timings, as text. */
int intp_batch_tuning(intp_batch b, char* buf, int64_t capacity);
int intp_batch_download_ctrl(intp_batch b, int32_t spline, double* host, int64_t count);

/* ---- serving: a request loop over resident splines ------------------ */
/* No reference counterpart (the reference builds and evaluates in the
 * caller's thread; InterpolationFunctionTemplate reuse,
 * InterpolationTemplate.hpp:118-143, is the closest analogue).  A server
 * owns the splines it builds.  A dispatcher thread admits every queued
 * request on each turn of its loop while earlier ones are still running
 * (continuous batching), spreads them over several HIP streams, and
 * records completion with HIP events:
 *   build: admitted when its control points fit the HBM budget next to
 *     the resident splines; least-recently-used splines with no pending
 *     evaluation are evicted to make room; a build that fits only after
 *     running work completes waits in the queue (evaluations of other
 *     splines pass it, its own wait with it);
 *   eval: runs on the spline's table once its build has finished (stream
 *     ordered, so an eval may be submitted right after its build).
 * Results are those of intp_build / intp_build_refined / intp_eval on the
 * same inputs, whatever else runs beside them.  Buffers passed to a
 * request (values, coords, out: device memory) must stay valid until it
 * completes.                                                             */
typedef struct intp_server_s* intp_server;

typedef struct {
    int64_t hbm_budget;   /* bytes for resident control points; 0: free HBM
                             at creation minus 8 GiB (at least 1 GiB)      */
    int32_t streams;      /* HIP streams the requests are spread over
                             (1..16; 0: 4)                                  */
    int32_t max_inflight; /* launched, unfinished requests (0: 256)        */
} intp_server_config;

#define INTP_REQ_BUILD 0
#define INTP_REQ_EVAL 1

/* request states */
#define INTP_REQ_QUEUED 0
#define INTP_REQ_RUNNING 1
#define INTP_REQ_DONE 2
#define INTP_REQ_FAILED 3

/* build precision */
#define INTP_PREC_F64 0
#define INTP_PREC_F32_REFINED 1

typedef struct {
    int64_t completed;
    int64_t failed;
    double bytes;           /* algorithmic bytes of the completed requests */
    double latency_p50_ms;  /* submit to completion */
    double latency_p99_ms;
    double latency_mean_ms;
    double latency_max_ms;
} intp_serve_type_counters;

typedef struct {
    intp_serve_type_counters type[2]; /* [INTP_REQ_BUILD], [INTP_REQ_EVAL] */
    int64_t resident_splines;
    int64_t resident_bytes;
    int64_t evictions;
    int64_t admission_waits;       /* turns a build waited for memory      */
    double first_submit_s;         /* server clock (s) of the first submit */
    double last_done_s;            /* and of the last completion, since the
                                      last reset                           */
} intp_serve_counters;

int intp_server_create(intp_ctx ctx, const intp_server_config* cfg, intp_server* out);
/* Waits for every submitted request, then frees the server's splines. */
int intp_server_destroy(intp_server s);
/* Queues a build of `values` (device, C-order) on grid `desc`; precision
 * INTP_PREC_F64 or INTP_PREC_F32_REFINED (with `sweeps` refinement
 * sweeps).  *request receives the request id, which is also the id of the
 * spline it builds.                                                      */
int intp_serve_build(intp_server s, const intp_grid_desc* desc, const double* values, int32_t precision,
                     int32_t sweeps, int64_t* request);
/* Queues an evaluation of spline `spline` (a build request id) with
 * intp_eval's arguments (device coords and out; orders is copied).       */
int intp_serve_eval(intp_server s, int64_t spline, int64_t npts, const double* coords, int64_t cs_dim,
                    int64_t cs_pt, int32_t nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt,
                    uint32_t flags, int64_t* request);
/* State (INTP_REQ_*) and, once finished, the intp_status of a request. */
int intp_serve_poll(intp_server s, int64_t request, int32_t* state, int32_t* status);
/* Blocks until the request is done or failed; returns its status.       */
int intp_serve_wait(intp_server s, int64_t request);
/* Blocks until every request submitted so far has finished.             */
int intp_serve_drain(intp_server s);
/* The client needs the spline no more: it is freed once idle.           */
int intp_serve_release(intp_server s, int64_t spline);
/* A new reference to a resident, built spline (release it with
 * intp_spline_release).                                                  */
int intp_serve_spline(intp_server s, int64_t spline, intp_spline* out);
/* Counters and latency percentiles of the requests finished since the
 * last reset (reset != 0 starts a new period after reading).             */
int intp_serve_stats(intp_server s, int32_t reset, intp_serve_counters* out);

/* Durations (ms) of the calls made with INTP_EVAL_TIMED on this thread
 * since the last query, oldest first: group_ms[i] covers the grouping
 * pass (0 when none ran), kernel_ms[i] the evaluation kernels alone, each
 * measured with HIP events on the call's own stream.  Writes up to
 * `capacity` entries, sets *count to the number of timed calls and
 * forgets them; waits for the last timed call to finish.  No reference
 * counterpart (a measurement hook for benchmarks and profiling).         */
int intp_eval_timings(double* group_ms, double* kernel_ms, int64_t capacity, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif
