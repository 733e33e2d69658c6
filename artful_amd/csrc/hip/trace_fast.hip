/* artful-hip — the tracers that may fuse multiply-adds (built with -ffp-contract=fast): FP32 fast and
 * mixed FP32/FP64 here, scheduled for instruction-level parallelism as the exact tracer is (mixed 4.5%
 * faster, FP32 fast 1.3%, the same bits); FP64 in trace_fp64.hip, on the default schedule. */
#include "queue_kernel.h"

namespace artf {

hipError_t launchTraceFast(const ArtfHipTraceArgs& a, hipStream_t stream) {
    switch (a.precision) {
    case ARTF_PRECISION_FP32_FAST:
        return launchTraceWith(artfHipDragInRange(a.frame.spin[0]) ? traceKernel<ARTF_PRECISION_FP32_FAST, true> : traceKernel<ARTF_PRECISION_FP32_FAST, false>, a, stream);
    case ARTF_PRECISION_MIXED:
        return launchTraceWith(artfHipDragInRange(a.frame.spin[0]) ? traceKernel<ARTF_PRECISION_MIXED, true> : traceKernel<ARTF_PRECISION_MIXED, false>, a, stream);
    case ARTF_PRECISION_FP64: return launchTrace64(a, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launchViewsFast(const ArtfHipBatchArgs& a, hipStream_t stream) {
    switch (a.precision) {
    case ARTF_PRECISION_FP32_FAST: return launchQueueWith<ARTF_PRECISION_FP32_FAST>(a, stream);
    case ARTF_PRECISION_MIXED: return launchQueueWith<ARTF_PRECISION_MIXED>(a, stream);
    case ARTF_PRECISION_FP64: return launchViews64(a, stream);
    default: return hipErrorInvalidValue;
    }
}

int viewsOccupancyFast(int precision) {
    switch (precision) {
    case ARTF_PRECISION_FP32_FAST: return queueOccupancy<ARTF_PRECISION_FP32_FAST>();
    case ARTF_PRECISION_MIXED: return queueOccupancy<ARTF_PRECISION_MIXED>();
    case ARTF_PRECISION_FP64: return viewsOccupancy64();
    default: return 0;
    }
}

} // namespace artf
