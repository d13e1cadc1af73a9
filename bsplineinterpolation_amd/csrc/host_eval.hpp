// host_eval.hpp -- evaluation on the calling host thread from a host copy
// of a spline's control points: the scalar members of the C++ drop-in
// (one point per call, where a device round trip would cost tens of
// microseconds) and host evaluation proxies.  Compiled with g++ without
// FMA contraction (host_eval.cpp).
#pragma once

#include <cstdint>

#include "setup.hpp"

namespace intp_amd {

constexpr int kHostMaxDim = 6;

// One spline as the host evaluator sees it: per-dimension setups (knots,
// range, periodicity), the order and the C-order control points.
struct HostSpline {
    const DimSetup* dims[kHostMaxDim];
    int dim = 0;
    int order = 0;
    const double* ctrl = nullptr;
    // refined splines (1 to 3 dimensions): the gradient table's backward
    // differences, diff[d * total + e] = c[e] - c[e - e_d] formed before
    // rounding; first partials use them in place of differences of ctrl
    const double* diff = nullptr;
    int64_t total = 0;
};

// Values / derivatives of npts points: coordinate d of point i at
// coords[d * cs_dim + i * cs_pt], output o at out[o * os_out + i * os_pt],
// derivative orders orders[o * dim + d].  With `bounds`, a clamped
// coordinate outside range(d) stops the call: returns the index of the
// first such point (outputs of the points before it are written), else -1.
int64_t host_eval_points(const HostSpline& s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                         int nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt, bool bounds);

// An evaluation proxy's per-point data (the reference's pre_calc_coef,
// BSpline.hpp:244-297): per dimension, the first stencil index mu - order
// (before periodic wrap) and the order + 1 basis values.
//   first[i * dim + d], weights[(i * dim + d) * (order + 1) + m]
// With `bounds`, returns the first point outside a clamped range (no
// data for it or later points), else -1.
int64_t host_proxy_weights(const HostSpline& s, int64_t npts, const double* coords, int64_t cs_dim, int64_t cs_pt,
                           bool bounds, int32_t* first, double* weights);

// Applies proxy data to the control points of a spline on the same grid:
// out[i * os_pt] = sum over the stencil of the weight products times the
// control points.
void host_proxy_apply(const HostSpline& s, int64_t npts, const int32_t* first, const double* weights, double* out,
                      int64_t os_pt);

}  // namespace intp_amd
