// eval_f32_d4.hip -- evaluation kernels for 4D splines with float tables.
#include "eval_impl.hpp"

namespace intp_amd {

hipError_t eval_dispatch_f32_d4(int order, bool grad, bool uniform, const EvalParams& p, hipStream_t s) {
    return eval_dispatch<4, float>(order, grad, uniform, p, s);
}

}  // namespace intp_amd
