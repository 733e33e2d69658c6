/* artful-hip — the FP32 tracer that reproduces the reference's arithmetic: this file is built with
 * -ffp-contract=off, so the compiler fuses no multiply-add and each rounding is the reference's. The only
 * fused ones are written out (scaledPlus, tracer.h): an exact product by a power of two added with one rounding,
 * where the reference's two roundings give the same bits. */
#include "queue_kernel.h"

namespace artf {

template __global__ void traceKernel<ARTF_PRECISION_FP32, false>(ArtfHipTraceArgs);
template __global__ void traceKernel<ARTF_PRECISION_FP32, true>(ArtfHipTraceArgs);

hipError_t launchTraceExact(const ArtfHipTraceArgs& a, hipStream_t stream) {
    return launchTraceWith(artfHipDragInRange(a.frame.spin[0]) ? traceKernel<ARTF_PRECISION_FP32, true> : traceKernel<ARTF_PRECISION_FP32, false>, a, stream);
}

hipError_t launchViewsExact(const ArtfHipBatchArgs& a, hipStream_t stream) { return launchQueueWith<ARTF_PRECISION_FP32>(a, stream); }

int viewsOccupancyExact() { return queueOccupancy<ARTF_PRECISION_FP32>(); }

} // namespace artf
