/* artful-hip — the FP32 tracer that reproduces the reference's arithmetic: this file is built with
 * -ffp-contract=off, so no multiply-add is fused and each rounding is the reference's. */
#include "queue_kernel.h"

namespace artf {

template __global__ void traceKernel<ARTF_PRECISION_FP32>(ArtfHipTraceArgs);

hipError_t launchTraceExact(const ArtfHipTraceArgs& a, hipStream_t stream) {
    return launchTraceWith(traceKernel<ARTF_PRECISION_FP32>, a, stream);
}

hipError_t launchViewsExact(const ArtfHipBatchArgs& a, hipStream_t stream) { return launchQueueWith<ARTF_PRECISION_FP32>(a, stream); }

int viewsOccupancyExact() { return queueOccupancy<ARTF_PRECISION_FP32>(); }

} // namespace artf
