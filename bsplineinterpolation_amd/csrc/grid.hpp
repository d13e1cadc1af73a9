// grid.hpp -- kernel parameters of one grid dimension from its host setup
// (shared by the single-spline and the batched entry points).
#pragma once

#include "kernels.hpp"
#include "setup.hpp"

namespace intp_amd {

// `stride`: elements between neighbouring control points along the
// dimension; the table pointers (knots, reciprocals, polynomials) are the
// caller's device copies.
inline DimParams make_dim_params(const DimSetup& s, int order, int64_t stride, const double* knots,
                                 const double* recip, const double* poly) {
    DimParams p{};
    p.knots = knots;
    p.recip = recip;
    p.n = s.n;
    p.stride = stride;
    p.nt = static_cast<int32_t>(s.knots.size());
    p.periodic = s.periodic ? 1 : 0;
    p.uniform = s.uniform ? 1 : 0;
    p.lo = s.lo;
    p.hi = s.hi;
    p.period = s.hi - s.lo;
    p.inv_dx = s.uniform ? 1.0 / s.dx : 0.0;
    p.shift = s.periodic ? 1.0 : 0.5 * static_cast<double>(order + 1);
    p.xmin = s.xmin;
    p.xmax = s.xmax;
    p.dx = s.dx;
    p.half = s.half;
    p.poly = poly;
    p.nleft = s.nleft;
    p.nright = s.nright;
    p.nclass = s.nclass;
    return p;
}

}  // namespace intp_amd
