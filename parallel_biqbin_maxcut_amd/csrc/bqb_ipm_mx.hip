// bqb_ipm_mx.hip -- launches of the batched interior-point kernel with the
// two products on the bf16 matrix cores (bqb_ipm_kernel.h, MIX = true);
// compiled apart from the FP64 instantiations to keep builds parallel.
#include <hip/hip_runtime.h>

#include "bqb_ipm_kernel.h"

namespace bqb {

template <int P, int R>
static hipError_t launch_ipm(const IpmBatchArgs &a, int nodes, hipStream_t s)
{
    hipLaunchKernelGGL((ipm_batch_kernel<P, R, true>), dim3(nodes), dim3(P * P), 0, s, a);
    return hipGetLastError();
}

}  // namespace bqb

using namespace bqb;

extern "C" hipError_t bqb_launch_ipm_mixed(const IpmBatchArgs *a, int nodes, int tile, hipStream_t s)
{
    switch (tile) {
    case 1604: return launch_ipm<16, 4>(*a, nodes, s);
    case 1605: return launch_ipm<16, 5>(*a, nodes, s);
    case 1606: return launch_ipm<16, 6>(*a, nodes, s);
    case 1607: return launch_ipm<16, 7>(*a, nodes, s);
    case 3204: return launch_ipm<32, 4>(*a, nodes, s);
    case 3205: return launch_ipm<32, 5>(*a, nodes, s);
    case 3206: return launch_ipm<32, 6>(*a, nodes, s);
    case 3207: return launch_ipm<32, 7>(*a, nodes, s);
    case 3208: return launch_ipm<32, 8>(*a, nodes, s);
    default: return hipErrorInvalidValue;
    }
}
