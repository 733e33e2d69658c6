/* artful-hip — the FP64 tracers (built with -ffp-contract=fast, as trace_fast.hip, and with the
 * compiler's default schedule: the instruction-level-parallelism schedule that trace_fast.hip's
 * tracers take made this one 3% slower). */
#include "queue_kernel.h"

namespace artf {

hipError_t launchTrace64(const ArtfHipTraceArgs& a, hipStream_t stream) { return launchTraceWith(traceKernel<ARTF_PRECISION_FP64>, a, stream); }

hipError_t launchViews64(const ArtfHipBatchArgs& a, hipStream_t stream) { return launchQueueWith<ARTF_PRECISION_FP64>(a, stream); }

int viewsOccupancy64() { return queueOccupancy<ARTF_PRECISION_FP64>(); }

} // namespace artf
