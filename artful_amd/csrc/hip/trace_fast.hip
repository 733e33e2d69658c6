/* artful-hip — the tracers that may fuse multiply-adds (built with -ffp-contract=fast): FP32 fast,
 * mixed FP32/FP64, and FP64. */
#include "queue_kernel.h"

namespace artf {

hipError_t launchTraceFast(const ArtfHipTraceArgs& a, hipStream_t stream) {
    switch (a.precision) {
    case ARTF_PRECISION_FP32_FAST: return launchTraceWith(traceKernel<ARTF_PRECISION_FP32_FAST>, a, stream);
    case ARTF_PRECISION_MIXED: return launchTraceWith(traceKernel<ARTF_PRECISION_MIXED>, a, stream);
    case ARTF_PRECISION_FP64: return launchTraceWith(traceKernel<ARTF_PRECISION_FP64>, a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launchViewsFast(const ArtfHipBatchArgs& a, hipStream_t stream) {
    switch (a.precision) {
    case ARTF_PRECISION_FP32_FAST: return launchQueueWith<ARTF_PRECISION_FP32_FAST>(a, stream);
    case ARTF_PRECISION_MIXED: return launchQueueWith<ARTF_PRECISION_MIXED>(a, stream);
    case ARTF_PRECISION_FP64: return launchQueueWith<ARTF_PRECISION_FP64>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

int viewsOccupancyFast(int precision) {
    switch (precision) {
    case ARTF_PRECISION_FP32_FAST: return queueOccupancy<ARTF_PRECISION_FP32_FAST>();
    case ARTF_PRECISION_MIXED: return queueOccupancy<ARTF_PRECISION_MIXED>();
    case ARTF_PRECISION_FP64: return queueOccupancy<ARTF_PRECISION_FP64>();
    default: return 0;
    }
}

} // namespace artf
