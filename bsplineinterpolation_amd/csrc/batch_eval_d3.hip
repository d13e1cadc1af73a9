// batch_eval_d3.hip -- batched evaluation kernels for 3D splines.
#define INTP_BATCH_EVAL_KERNELS
#include "eval_impl.hpp"
#include "batch_impl.hpp"

namespace intp_amd {

hipError_t batch_eval_dispatch_d3(int order, const BatchEvalArgs& a, int blocks, int threads, hipStream_t s) {
    return batch_eval_dispatch<3>(order, a, blocks, threads, s);
}

}  // namespace intp_amd
