// eval_f64_d5.hip -- evaluation kernels for 5D splines with double tables.
#include "eval_impl.hpp"

namespace intp_amd {

hipError_t eval_dispatch_f64_d5(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    return eval_dispatch<5, double>(order, grad, uniform, p, s);
}

}  // namespace intp_amd
