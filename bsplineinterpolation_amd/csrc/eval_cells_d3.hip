// eval_cells_d3.hip -- gradient kernels for 3D refined builds on their
// gradient table (value and first partials from the FP32 cells).
#include "eval_impl.hpp"

namespace intp_amd {

hipError_t eval_cells_d3(int order, bool uniform, const EvalParams& p, hipStream_t s) {
    return eval_dispatch<3, Cell3>(order, true, uniform, p, s);
}

}  // namespace intp_amd
