/*
 * intp_oracle.c -- CPU restatement of the reference algorithm; TEST
 * INFRASTRUCTURE ONLY (see intp_oracle.h).  Plain C11 + pthreads.
 *
 * Each function names the reference lines it follows.  The arithmetic order
 * of the reference is kept where it decides rounding (basis recursion, LU,
 * substitution, stencil sums), so that the oracle agrees with the compiled
 * reference to round-off (checked in tests/test_oracle_vs_ref.py).
 */
#include "intp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }

/* ===================================================================== */
/* knots                                                                 */
/* ===================================================================== */

/* Knot count: Interpolation.hpp:338-340 (uniform) and :394-395
 * (non-uniform), with INTP_PERIODIC_NO_DUMMY_POINT (:330-334).           */
int64_t orc_knots_count(int order, int periodic, int64_t n_stored) {
    const int64_t np = n_stored + (periodic ? 1 : 0);
    const int64_t extra =
        periodic ? 2 * order + (1 - order % 2) : (int64_t)order + 1;
    return np + extra;
}

/* Uniform knot vector: Interpolation.hpp:322-362. */
void orc_knots_uniform(int order, int periodic, int64_t n_stored, double xmin,
                       double xmax, double* xs, double* dx_out) {
    const int64_t n = n_stored + (periodic ? 1 : 0);
    const double dx = (xmax - xmin) / (double)(n - 1);
    const int64_t extra =
        periodic ? 2 * order + (1 - order % 2) : (int64_t)order + 1;
    const int64_t size = n + extra;
    for (int64_t i = 0; i < size; ++i) xs[i] = xmin;
    if (periodic) {
        for (int64_t i = 0; i < size; ++i)
            xs[i] = xmin + ((double)i - 0.5 * (double)extra) * dx;
    } else {
        for (int64_t i = order + 1; i < size - order - 1; ++i)
            xs[i] = xmin + ((double)i - 0.5 * (double)extra) * dx;
        for (int64_t i = size - order - 1; i < size; ++i) xs[i] = xmax;
    }
    *dx_out = dx;
}

/* Non-uniform knot vector: Interpolation.hpp:365-447.  xs holds
 * n_stored + periodic node coordinates.  Returns -1 if the coordinates are
 * not strictly increasing (:448-454) or order 0 is requested (the
 * reference divides the moving-window sum by the order, :439).          */
int orc_knots_nonuniform(int order, int periodic, int64_t n_stored,
                         const double* x, double* xs) {
    const int64_t n = n_stored + (periodic ? 1 : 0);
    const int64_t k = order;
    const int64_t size = periodic ? n + 2 * k + (1 - k % 2) : n + k + 1;
    for (int64_t i = 0; i + 1 < n; ++i)
        if (!(x[i + 1] > x[i])) return -1;
    if (periodic) {
        /* :401-417 */
        for (int64_t i = k + 1; i < k + n; ++i) {
            const double prev = x[i - k - 1], present = x[i - k];
            xs[i] = (k % 2 == 0) ? 0.5 * (prev + present) : present;
        }
        const double period = x[n - 1] - x[0];
        for (int64_t i = 0; i < k + 1; ++i) {
            xs[i] = xs[n + i - 1] - period;
            xs[size - i - 1] = xs[size - i - n] + period;
        }
    } else {
        if (k == 0) return -1;
        /* :418-447 */
        for (int64_t i = 0; i < k + 1; ++i) xs[i] = x[0];
        double window_sum = 0.0;
        for (int64_t i = 1; i < k; ++i) window_sum += x[i];
        for (int64_t i = k + 1; i < n; ++i) {
            window_sum += x[i - 1];
            xs[i] = window_sum / (double)k;
            window_sum -= x[i - k];
        }
        for (int64_t i = n; i < n + k + 1; ++i) xs[i] = x[n - 1];
    }
    return 0;
}

/* ===================================================================== */
/* basis and span                                                        */
/* ===================================================================== */

/* Cox-de Boor triangle, right-aligned for reduced orders:
 * BSpline.hpp:83-111.  t[mu] is the left knot of the span.              */
void orc_base_spline_value(const double* t, int64_t mu, double x, int order,
                           int spline_order, double* b) {
    const int k = order;
    for (int i = 0; i <= k; ++i) b[i] = 0.0;
    b[k] = 1.0;
    for (int i = 1; i <= spline_order; ++i) {
        const int idx_begin = k - i;
        for (int j = 0; j <= i; ++j) {
            const int64_t left = mu - (i - j);
            const int64_t right = mu + j + 1;
            const double a =
                (j == 0) ? 0.0
                         : b[idx_begin + j] * (x - t[left]) /
                               (t[right - 1] - t[left]);
            const double c = (idx_begin + j == k)
                                 ? 0.0
                                 : b[idx_begin + j + 1] * (t[right] - x) /
                                       (t[right] - t[left + 1]);
            b[idx_begin + j] = a + c;
        }
    }
}

/* Uniform-grid span hint: Interpolation.hpp:277-295 (and :306-318, the
 * same formula for derivatives).  Non-uniform dims use `order`.          */
int64_t orc_hint(const orc_spline* s, int d, double x) {
    const int k = s->order;
    if (!s->uniform[d]) return k;
    const double shift = s->periodic[d] ? 1.0 : 0.5 * (double)(k + 1);
    double v = (x - s->lo[d]) / s->dx[d] - shift;
    if (!(v > 0.0)) v = 0.0; /* std::max(0., v) returns 0 for NaN too */
    v = ceil(v);
    const int64_t cap = s->nt[d] - k - 2;
    /* the reference casts to size_t before std::min; clamp first so that
     * huge coordinates stay defined                                     */
    if (v >= (double)cap) return cap;
    return imin(cap, (int64_t)v + k);
}

/* Span search with periodic wrap: BSpline.hpp:125-157.  Modifies *x for
 * periodic dims exactly like the reference.  Returns mu in [k, last].    */
int64_t orc_find_span(const orc_spline* s, int d, double* xp, int64_t hint,
                      int64_t last) {
    const int k = s->order;
    const double* t = s->knots[d];
    if (s->periodic[d]) {
        const double lo = s->lo[d];
        const double period = s->hi[d] - lo;
        const double x = *xp;
        *xp = lo + fmod(x - lo, period) + (x < lo ? period : 0.0);
    }
    const double x = *xp;
    if (t[hint] <= x && t[hint + 1] > x) return hint;
    /* upper_bound over [t_{k+1}, t_{last+1}) then step back */
    int64_t lo_i = k + 1, hi_i = last + 1; /* search in [lo_i, hi_i) */
    while (lo_i < hi_i) {
        const int64_t mid = lo_i + (hi_i - lo_i) / 2;
        if (x < t[mid])
            hi_i = mid;
        else
            lo_i = mid + 1;
    }
    return lo_i - 1;
}

/* ===================================================================== */
/* band matrices and LU                                                  */
/* ===================================================================== */

int orc_band_alloc(orc_band* b, int64_t n, int64_t p, int64_t q,
                   int extended) {
    memset(b, 0, sizeof(*b));
    b->n = n;
    b->p = p;
    b->q = q;
    b->extended = extended;
    b->main = (double*)calloc((size_t)(n * (1 + p + q)) + 1, sizeof(double));
    if (extended) {
        b->right = (double*)calloc((size_t)((n - q - 1) * p) + 1,
                                   sizeof(double));
        b->bottom = (double*)calloc((size_t)((n - p - 1) * q) + 1,
                                    sizeof(double));
    }
    return b->main ? 0 : -1;
}

void orc_band_free(orc_band* b) {
    free(b->main);
    free(b->right);
    free(b->bottom);
    memset(b, 0, sizeof(*b));
}

/* BandMatrix.hpp:50-60 */
static double* main_at(const orc_band* b, int64_t i, int64_t j) {
    return &b->main[j * (1 + b->p + b->q) + (i + b->q - j)];
}
/* ExtendedBandMatrix side bands: BandMatrix.hpp:121-136 */
static double* side_at(const orc_band* b, int64_t i, int64_t j) {
    if (j > i + b->q) return &b->right[i * b->p + (j + b->p - b->n)];
    return &b->bottom[j * b->q + (i + b->q - b->n)];
}
/* ExtendedBandMatrix::operator(): BandMatrix.hpp:137-146 */
double* orc_band_at(orc_band* b, int64_t i, int64_t j) {
    if (b->extended && (j > i + b->q || i > j + b->p)) return side_at(b, i, j);
    return main_at(b, i, j);
}

/* No-pivot LU: BandLU.hpp:103-118 (band) and :159-213 (extended). */
void orc_band_lu(orc_band* b) {
    const int64_t n = b->n, p = b->p, q = b->q;
    if (!b->extended) {
        for (int64_t k = 0; k + 1 < n; ++k) {
            for (int64_t i = k + 1; i < imin(k + p + 1, n); ++i)
                *main_at(b, i, k) /= *main_at(b, k, k);
            for (int64_t j = k + 1; j < imin(k + q + 1, n); ++j)
                for (int64_t i = k + 1; i < imin(k + p + 1, n); ++i)
                    *main_at(b, i, j) -= *main_at(b, i, k) * *main_at(b, k, j);
        }
        return;
    }
    for (int64_t k = 0; k + 1 < n; ++k) {
        for (int64_t i = k + 1; i < imin(k + p + 1, n); ++i)
            *main_at(b, i, k) /= *main_at(b, k, k);
        for (int64_t i = imax(n - q, k + p + 1); i < n; ++i)
            *side_at(b, i, k) /= *main_at(b, k, k);
        for (int64_t j = k + 1; j < imin(k + q + 1, n); ++j)
            for (int64_t i = k + 1; i < imin(k + p + 1, n); ++i)
                *main_at(b, i, j) -= *main_at(b, i, k) * *main_at(b, k, j);
        for (int64_t i = k + 1; i < imin(k + p + 1, n); ++i)
            for (int64_t j = imax(n - p, k + q + 1); j < n; ++j)
                *orc_band_at(b, i, j) -= *main_at(b, i, k) * *side_at(b, k, j);
        for (int64_t j = k + 1; j < imin(k + q + 1, n); ++j)
            for (int64_t i = imax(n - q, k + p + 1); i < n; ++i)
                *orc_band_at(b, i, j) -= *side_at(b, i, k) * *main_at(b, k, j);
        if (k < imax(n - p - 1, n - q - 1))
            for (int64_t i = imax(n - q, k + p + 1); i < n; ++i)
                for (int64_t j = imax(n - p, k + q + 1); j < n; ++j)
                    *main_at(b, i, j) -= *side_at(b, i, k) * *side_at(b, k, j);
    }
}

/* In-place substitution on one strided line: BandLU.hpp:121-143 (band) and
 * :216-259 (extended), same loop order.                                   */
void orc_band_solve(const orc_band* b, double* x, int64_t s) {
    const int64_t n = b->n, p = b->p, q = b->q;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = j + 1; i < imin(j + p + 1, n); ++i)
            x[i * s] -= *main_at(b, i, j) * x[j * s];
        if (b->extended && j < n - p - 1)
            for (int64_t i = imax(n - q, j + p + 1); i < n; ++i)
                x[i * s] -= *side_at(b, i, j) * x[j * s];
    }
    for (int64_t j = n - 1; j >= 0; --j) {
        x[j * s] /= *main_at(b, j, j);
        for (int64_t i = (j < q ? 0 : j - q); i < j; ++i)
            x[i * s] -= *main_at(b, i, j) * x[j * s];
        if (b->extended && j > n - p - 1)
            for (int64_t i = 0; i < j - q; ++i)
                x[i * s] -= *side_at(b, i, j) * x[j * s];
    }
}

/* ===================================================================== */
/* template: knots + collocation matrices + LU                           */
/* ===================================================================== */

static void set_range(orc_spline* s, int d) {
    /* BSpline.hpp:224-226 */
    const int k = s->order;
    s->lo[d] = s->knots[d][k];
    s->hi[d] = s->knots[d][s->nt[d] - k - (2 - k % 2)];
}

/* Collocation matrix per dimension: InterpolationTemplate.hpp:254-446
 * (INTP_PERIODIC_NO_DUMMY_POINT branch).                                  */
static int build_solver(orc_spline* s) {
    const int k = s->order;
    double bsv[ORC_MAXD][ORC_MAXK + 1];
    memset(bsv, 0, sizeof(bsv));
    for (int d = 0; d < s->dim; ++d) {
        if (s->periodic[d] && s->uniform[d]) {
            /* :273-280 */
            const double* t = s->knots[d];
            orc_base_spline_value(t, k, t[k] + (double)(1 - k % 2) * s->dx[d] * 0.5,
                                  k, k, bsv[d]);
        }
    }
    for (int d = 0; d < s->dim; ++d) {
        const int periodic = s->periodic[d], uniform = s->uniform[d];
        const int64_t n = s->n[d];
        const int64_t bw = periodic ? k / 2 : (k == 0 ? 0 : k - 1);
        orc_band* m = &s->lu[d];
        if (orc_band_alloc(m, n, bw, bw, periodic)) return -1;
        const double* t = s->knots[d];
        const int64_t nt = s->nt[d];
        for (int64_t i = 0; i < n; ++i) {
            if (!periodic && (i == 0 || i == n - 1)) {
                *orc_band_at(m, i, i) = 1.0;
                continue;
            }
            int64_t knot_ind;
            const int is_internal = i > k / 2 && i < n - k / 2 - 1;
            if (uniform) {
                knot_ind = periodic ? i + k
                                    : imin(nt - k - 2, i > k / 2 ? i + (k + 1) / 2 : k);
                if (!periodic &&
                    (knot_ind <= 2 * k + 1 || knot_ind >= nt - 2 * k - 2)) {
                    const double x = s->lo[d] + (double)i * s->dx[d];
                    orc_base_spline_value(t, knot_ind, x, k, k, bsv[d]);
                }
            } else {
                double x = s->xin[d][i];
                int64_t it;
                if (periodic)
                    it = i + k;
                else if (i == 0)
                    it = k;
                else if (i == s->nxin[d] - 1)
                    it = nt - (k + 2);
                else
                    it = orc_find_span(s, d, &x, i + 1, imin(nt - k - 1, i + k));
                knot_ind = it;
                orc_base_spline_value(t, it, x, k, k, bsv[d]);
            }
            const int64_t s_num = periodic ? (k | 1)
                                  : k == 1 ? 1
                                  : (uniform && is_internal) ? (k | 1)
                                                             : k + 1;
            for (int64_t j = 0; j < s_num; ++j) {
                const int64_t row = (i + (periodic ? bw : 0)) % n;
                const int64_t col = (knot_ind - k + j) % n;
                *orc_band_at(m, row, col) = bsv[d][j];
            }
        }
        orc_band_lu(m);
    }
    return 0;
}

orc_spline* orc_new(void) { return (orc_spline*)calloc(1, sizeof(orc_spline)); }
void orc_delete(orc_spline* s) {
    if (!s) return;
    orc_free(s);
    free(s);
}

void orc_free(orc_spline* s) {
    for (int d = 0; d < ORC_MAXD; ++d) {
        free(s->knots[d]);
        free(s->xin[d]);
        if (s->lu[d].main) orc_band_free(&s->lu[d]);
    }
    free(s->ctrl);
    memset(s, 0, sizeof(*s));
}

/* InterpolationFunctionTemplate ctor: InterpolationTemplate.hpp:60-79 and
 * the partial InterpolationFunction ctor, Interpolation.hpp:111-121.      */
int orc_template_init(orc_spline* s, int dim, int order, const int* periodic,
                      const int64_t* n, const double* xmin,
                      const double* xmax, const double* const* coords) {
    memset(s, 0, sizeof(*s));
    if (dim < 1 || dim > ORC_MAXD || order < 0 || order > ORC_MAXK) return -1;
    s->dim = dim;
    s->order = order;
    s->nctrl = 1;
    for (int d = 0; d < dim; ++d) {
        s->periodic[d] = periodic ? periodic[d] != 0 : 0;
        s->n[d] = n[d];
        s->nctrl *= n[d];
        s->nt[d] = orc_knots_count(order, s->periodic[d], n[d]);
        s->knots[d] = (double*)calloc((size_t)s->nt[d], sizeof(double));
        if (coords && coords[d]) {
            s->uniform[d] = 0;
            s->nxin[d] = n[d] + s->periodic[d];
            s->xin[d] = (double*)malloc((size_t)s->nxin[d] * sizeof(double));
            memcpy(s->xin[d], coords[d], (size_t)s->nxin[d] * sizeof(double));
            if (orc_knots_nonuniform(order, s->periodic[d], n[d], coords[d],
                                     s->knots[d])) {
                orc_free(s);
                return -1;
            }
            s->dx[d] = 0.0;
        } else {
            s->uniform[d] = 1;
            orc_knots_uniform(order, s->periodic[d], n[d], xmin[d], xmax[d],
                              s->knots[d], &s->dx[d]);
        }
        set_range(s, d);
    }
    if (build_solver(s)) {
        orc_free(s);
        return -1;
    }
    return 0;
}

int orc_bspline_init(orc_spline* s, int dim, int order, const int* periodic,
                     const int64_t* n, const double* const* knots,
                     const int64_t* nknots, const double* ctrl) {
    memset(s, 0, sizeof(*s));
    s->dim = dim;
    s->order = order;
    s->nctrl = 1;
    for (int d = 0; d < dim; ++d) {
        s->periodic[d] = periodic ? periodic[d] != 0 : 0;
        s->uniform[d] = 0; /* explicit knots: hint = order */
        s->n[d] = n[d];
        s->nctrl *= n[d];
        s->nt[d] = nknots[d];
        s->knots[d] = (double*)malloc((size_t)nknots[d] * sizeof(double));
        memcpy(s->knots[d], knots[d], (size_t)nknots[d] * sizeof(double));
        set_range(s, d);
    }
    s->ctrl = (double*)malloc((size_t)s->nctrl * sizeof(double));
    memcpy(s->ctrl, ctrl, (size_t)s->nctrl * sizeof(double));
    return 0;
}

/* ===================================================================== */
/* control-point solve (InterpolationTemplate.hpp:448-580)               */
/* ===================================================================== */

typedef struct {
    const orc_spline* s;
    int axis;
    int64_t outer, inner, n;
    int64_t nlines;
    int64_t block, nblocks;
    int64_t next; /* next block (guarded by mu) */
    pthread_mutex_t mu;
} solve_job;

static void solve_lines(solve_job* job, int64_t begin, int64_t end) {
    const orc_band* b = &job->s->lu[job->axis];
    for (int64_t l = begin; l < end; ++l) {
        const int64_t o = l / job->inner, i = l % job->inner;
        orc_band_solve(b, job->s->ctrl + o * job->n * job->inner + i,
                       job->inner);
    }
}

static void* solve_worker(void* arg) {
    solve_job* job = (solve_job*)arg;
    for (;;) {
        pthread_mutex_lock(&job->mu);
        const int64_t blk = job->next++;
        pthread_mutex_unlock(&job->mu);
        if (blk >= job->nblocks) break;
        solve_lines(job, blk * job->block, (blk + 1) * job->block);
    }
    return NULL;
}

/* Copies f into the weights with the periodic shift +order/2 (:452-462),
 * then solves every line of every axis, last axis first (the reference
 * solves the contiguous axis of an array rotated after each pass,
 * :501-573; the per-line arithmetic is identical, so is the result).
 * The line set of an axis is split into floor(sqrt(lines)) blocks run on
 * nthreads workers plus the calling thread (:547-572).                    */
int orc_solve(orc_spline* s, const double* f, int nthreads) {
    const int D = s->dim, k = s->order;
    free(s->ctrl);
    s->ctrl = (double*)malloc((size_t)s->nctrl * sizeof(double));
    if (!s->ctrl) return -1;
    int64_t idx[ORC_MAXD];
    for (int d = 0; d < D; ++d) idx[d] = 0;
    for (int64_t lin = 0; lin < s->nctrl; ++lin) {
        int64_t w = 0;
        for (int d = 0; d < D; ++d) {
            int64_t id = idx[d];
            if (s->periodic[d]) id = (id + s->n[d] + k / 2) % s->n[d];
            w = w * s->n[d] + id;
        }
        s->ctrl[w] = f[lin];
        for (int d = D - 1; d >= 0; --d) {
            if (++idx[d] < s->n[d]) break;
            idx[d] = 0;
        }
    }
    for (int axis = D - 1; axis >= 0; --axis) {
        solve_job job;
        memset(&job, 0, sizeof(job));
        job.s = s;
        job.axis = axis;
        job.n = s->n[axis];
        job.outer = 1;
        job.inner = 1;
        for (int d = 0; d < axis; ++d) job.outer *= s->n[d];
        for (int d = axis + 1; d < D; ++d) job.inner *= s->n[d];
        job.nlines = job.outer * job.inner;
        if (nthreads <= 0) {
            solve_lines(&job, 0, job.nlines);
            continue;
        }
        job.nblocks = (int64_t)sqrt((double)job.nlines);
        job.block = job.nblocks == 0 ? job.nlines : job.nlines / job.nblocks;
        pthread_mutex_init(&job.mu, NULL);
        pthread_t th[64];
        const int nt = nthreads > 64 ? 64 : nthreads;
        for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, solve_worker, &job);
        solve_lines(&job, job.nblocks * job.block, job.nlines);
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
        pthread_mutex_destroy(&job.mu);
    }
    return 0;
}

/* ===================================================================== */
/* evaluation                                                            */
/* ===================================================================== */

static int64_t ipow(int64_t b, int e) {
    int64_t r = 1;
    while (e-- > 0) r *= b;
    return r;
}

/* Value: BSpline.hpp:305-335 (INTP_CELL_LAYOUT branch: the (k+1)^D stencil
 * is summed with dim 0 fastest, coef = c * b_0 * b_1 * ...), via
 * Interpolation.hpp:132-146,297-300 for the hints.  The cell table is a
 * re-indexing of the plain table (BSpline.hpp:677-731); the index is
 * computed here directly, wrapping periodic dims.                        */
double orc_eval(const orc_spline* s, const double* xin) {
    const int D = s->dim, k = s->order;
    double base[ORC_MAXD][ORC_MAXK + 1];
    int64_t mu[ORC_MAXD];
    for (int d = 0; d < D; ++d) {
        double x = xin[d];
        const int64_t hint = orc_hint(s, d, x);
        mu[d] = orc_find_span(s, d, &x, hint, s->nt[d] - k - 2);
        orc_base_spline_value(s->knots[d], mu[d], x, k, k, base[d]);
    }
    const int64_t nb = ipow(k + 1, D);
    double v = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        int64_t off = 0, comb = i;
        int64_t digit[ORC_MAXD];
        for (int d = 0; d < D; ++d) {
            digit[d] = comb % (k + 1);
            comb /= (k + 1);
        }
        for (int d = 0; d < D; ++d) {
            int64_t id = mu[d] - k + digit[d];
            if (s->periodic[d]) id %= s->n[d];
            off = off * s->n[d] + id;
        }
        double coef = s->ctrl[off];
        for (int d = 0; d < D; ++d) coef *= base[d][digit[d]];
        v += coef;
    }
    return v;
}

/* Mixed partial derivative: BSpline.hpp:393-532 (cell-layout gather,
 * :433-455), via Interpolation.hpp:177-205,302-319.                       */
double orc_derivative(const orc_spline* s, const double* xin,
                      const int64_t* r) {
    const int D = s->dim, k = s->order;
    int so[ORC_MAXD];
    for (int d = 0; d < D; ++d) {
        so[d] = (int64_t)k >= r[d] ? k - (int)r[d] : k + 1;
        if (so[d] > k) return 0.0;
    }
    double base[ORC_MAXD][ORC_MAXK + 1];
    int64_t mu[ORC_MAXD];
    for (int d = 0; d < D; ++d) {
        double x = xin[d];
        const int64_t hint = orc_hint(s, d, x);
        mu[d] = orc_find_span(s, d, &x, hint, s->nt[d] - k - 2);
        orc_base_spline_value(s->knots[d], mu[d], x, k, so[d], base[d]);
    }
    const int64_t nb = ipow(k + 1, D);
    double* lc = (double*)malloc((size_t)nb * sizeof(double));
    double* ls = (double*)malloc((size_t)nb * sizeof(double));
    for (int64_t i = 0; i < nb; ++i) {
        int64_t dig[ORC_MAXD], comb = i;
        for (int d = 0; d < D; ++d) {
            dig[d] = comb % (k + 1);
            comb /= (k + 1);
        }
        double coef = 1.0;
        for (int d = 0; d < D; ++d) coef *= base[d][dig[d]];
        int64_t off = 0, loc = 0;
        for (int d = 0; d < D; ++d) {
            int64_t id = mu[d] - k + dig[d];
            if (s->periodic[d]) id %= s->n[d];
            off = off * s->n[d] + id;
            loc = loc * (k + 1) + dig[d];
        }
        ls[loc] = coef;
        lc[loc] = s->ctrl[off];
    }
    /* differencing of local control points: :489-520 */
    for (int d = 0; d < D; ++d) {
        if (so[d] == k) continue;
        const int64_t stride = ipow(k + 1, D - 1 - d);
        const int64_t hs = nb / (k + 1);
        const double* t = s->knots[d];
        for (int64_t h = 0; h < hs; ++h) {
            /* h enumerates the other dims; base offset of the line */
            const int64_t hi = h / stride, lo = h % stride;
            double* it = lc + hi * stride * (k + 1) + lo;
            for (int kk = k; kk > so[d]; --kk) {
                for (int j = kk; j > 0; --j) {
                    const int64_t a = k + j - kk;
                    it[a * stride] = (double)kk * (it[a * stride] - it[(a - 1) * stride]) /
                                     (t[mu[d] + j] - t[mu[d] + j - kk]);
                }
            }
        }
    }
    double v = 0.0;
    for (int64_t i = 0; i < nb; ++i) v += ls[i] * lc[i];
    free(lc);
    free(ls);
    return v;
}

typedef struct {
    const orc_spline* s;
    int64_t begin, end, npts;
    const double* coords;
    int nderiv;
    const int64_t* orders;
    double* out;
    int bounds;
    int bad;
} eval_job;

static void* eval_worker(void* arg) {
    eval_job* j = (eval_job*)arg;
    const orc_spline* s = j->s;
    const int D = s->dim;
    for (int64_t p = j->begin; p < j->end; ++p) {
        const double* x = j->coords + p * D;
        if (j->bounds) {
            /* boundary_check_: Interpolation.hpp:482-491 */
            for (int d = 0; d < D; ++d)
                if (!s->periodic[d] && (x[d] < s->lo[d] || x[d] > s->hi[d])) j->bad = 1;
        }
        for (int o = 0; o < j->nderiv; ++o) {
            const int64_t* r = j->orders + o * D;
            int zero = 1;
            for (int d = 0; d < D; ++d) zero &= r[d] == 0;
            j->out[o * j->npts + p] = zero ? orc_eval(s, x) : orc_derivative(s, x, r);
        }
    }
    return NULL;
}

int orc_eval_batch(const orc_spline* s, int64_t npts, const double* coords,
                   int nderiv, const int64_t* orders, double* out,
                   int nthreads, int bounds) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    eval_job jobs[256];
    pthread_t th[256];
    const int64_t chunk = (npts + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].s = s;
        jobs[t].begin = imin(npts, t * chunk);
        jobs[t].end = imin(npts, (t + 1) * chunk);
        jobs[t].npts = npts;
        jobs[t].coords = coords;
        jobs[t].nderiv = nderiv;
        jobs[t].orders = orders;
        jobs[t].out = out;
        jobs[t].bounds = bounds;
        jobs[t].bad = 0;
    }
    if (nthreads == 1) {
        eval_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, eval_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    for (int t = 0; t < nthreads; ++t)
        if (jobs[t].bad) return -1;
    return 0;
}

int64_t orc_nt(const orc_spline* s, int d) { return s->nt[d]; }
const double* orc_knots_ptr(const orc_spline* s, int d) { return s->knots[d]; }
void orc_range(const orc_spline* s, int d, double* lo, double* hi) {
    *lo = s->lo[d];
    *hi = s->hi[d];
}
const double* orc_ctrl_ptr(const orc_spline* s) { return s->ctrl; }
