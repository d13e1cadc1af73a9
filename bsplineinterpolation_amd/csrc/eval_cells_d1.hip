// eval_cells_d1.hip -- gradient kernels for 1D refined builds on their
// gradient table (value and first partials from the FP32 cells).
#include "eval_impl.hpp"

namespace intp_amd {

hipError_t eval_cells_d1(int order, bool uniform, const EvalParams& p, hipStream_t s) {
    return eval_dispatch<1, Cell1>(order, true, uniform, p, s);
}

}  // namespace intp_amd
