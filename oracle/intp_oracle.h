/*
 * intp_oracle.h -- CPU restatement of the reference B-spline interpolation
 * path (12ff54e/BSplineInterpolation), used ONLY as test infrastructure:
 * the tests in tests/, the reference leg of smoke(), and the cpu_baseline
 * leg of bench.py.  Nothing in the product (bsplineinterpolation_amd/,
 * include/) links or calls it.
 *
 * Conventions follow the reference test build (test/CMakeLists.txt:46-47):
 * INTP_PERIODIC_NO_DUMMY_POINT (periodic dims store n values, dx = P/n) and
 * INTP_CELL_LAYOUT (value evaluation sums with dim 0 fastest).  Every
 * function cites the reference file:line it restates.
 */
#ifndef INTP_ORACLE_H
#define INTP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXD 8
#define ORC_MAXK 15

/* Band storage of src/include/BandMatrix.hpp:12-16,107-136:
 *   main   : n x (1+p+q), A(i,j) = main[j*(1+p+q) + (i+q-j)]
 *   right  : (n-q-1) x p, A(i,j) = right[i*p + (j+p-n)]   for j > i+q
 *   bottom : (n-p-1) x q, A(i,j) = bottom[j*q + (i+q-n)]  for i > j+p   */
typedef struct {
    int64_t n, p, q;
    int extended;
    double* main;
    double* right;
    double* bottom;
} orc_band;

typedef struct {
    int dim, order;
    int periodic[ORC_MAXD];
    int uniform[ORC_MAXD];
    int64_t n[ORC_MAXD];  /* stored values (= control points) per dim   */
    int64_t nt[ORC_MAXD]; /* knots per dim                              */
    double* knots[ORC_MAXD];
    double lo[ORC_MAXD], hi[ORC_MAXD]; /* range_ (BSpline.hpp:224-226)   */
    double dx[ORC_MAXD];
    double* xin[ORC_MAXD]; /* input node coordinates (non-uniform dims) */
    int64_t nxin[ORC_MAXD];
    orc_band lu[ORC_MAXD];
    double* ctrl; /* control points, C-order over n[]                  */
    int64_t nctrl;
} orc_spline;

/* ---- knots (Interpolation.hpp:322-480, BSpline.hpp:217-227) ---------- */
int64_t orc_knots_count(int order, int periodic, int64_t n_stored);
void orc_knots_uniform(int order, int periodic, int64_t n_stored, double xmin,
                       double xmax, double* knots, double* dx);
int orc_knots_nonuniform(int order, int periodic, int64_t n_stored,
                         const double* xs, double* knots);

/* ---- basis and span (BSpline.hpp:83-171, Interpolation.hpp:277-295) --- */
void orc_base_spline_value(const double* knots, int64_t mu, double x,
                           int order, int spline_order, double* out);
int64_t orc_hint(const orc_spline* s, int d, double x);
int64_t orc_find_span(const orc_spline* s, int d, double* x, int64_t hint,
                      int64_t last);

/* ---- band LU (BandLU.hpp:103-259) ----------------------------------- */
int orc_band_alloc(orc_band* b, int64_t n, int64_t p, int64_t q, int extended);
void orc_band_free(orc_band* b);
double* orc_band_at(orc_band* b, int64_t i, int64_t j);
void orc_band_lu(orc_band* b);
void orc_band_solve(const orc_band* b, double* x, int64_t stride);

/* ---- interpolation function (InterpolationTemplate.hpp:60-580) ------- */
/* coords[d] == NULL -> uniform on [xmin[d], xmax[d]]; else non-uniform with
 * n[d] + periodic[d] node coordinates.  Returns 0 on success.             */
int orc_template_init(orc_spline* s, int dim, int order, const int* periodic,
                      const int64_t* n, const double* xmin,
                      const double* xmax, const double* const* coords);
int orc_solve(orc_spline* s, const double* f, int nthreads);
void orc_free(orc_spline* s);

/* explicit knots + control points (BSpline ctor, BSpline.hpp:188-210) */
int orc_bspline_init(orc_spline* s, int dim, int order, const int* periodic,
                     const int64_t* n, const double* const* knots,
                     const int64_t* nknots, const double* ctrl);

/* ---- evaluation (BSpline.hpp:305-368, 393-532) ----------------------- */
double orc_eval(const orc_spline* s, const double* x);
double orc_derivative(const orc_spline* s, const double* x,
                      const int64_t* orders);
/* coords AoS [npts][dim]; orders [nderiv][dim]; out [nderiv][npts].  An
 * all-zero order row is evaluated through orc_eval (operator()), others
 * through orc_derivative (derivative()).  bounds != 0 reproduces at() /
 * derivative_at(): returns -1 if any clamped coordinate is out of range.  */
int orc_eval_batch(const orc_spline* s, int64_t npts, const double* coords,
                   int nderiv, const int64_t* orders, double* out,
                   int nthreads, int bounds);

/* helpers for the Python side */
int64_t orc_nt(const orc_spline* s, int d);
const double* orc_knots_ptr(const orc_spline* s, int d);
void orc_range(const orc_spline* s, int d, double* lo, double* hi);
orc_spline* orc_new(void);
void orc_delete(orc_spline* s);
const double* orc_ctrl_ptr(const orc_spline* s);

#ifdef __cplusplus
}
#endif

#endif
