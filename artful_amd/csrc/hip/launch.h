/* artful-hip — what the kernel files hand each other: launchers, one per translation unit, so that
 * each unit can be built with its own floating-point contraction rule. */
#pragma once

#include <hip/hip_runtime.h>

#include "artful_hip.h"

namespace artf {

/* The tile of the tile tracer and the adjoint: 8 x 8 pixels, one wave a workgroup. */
constexpr int kTraceTile = 8;
constexpr int kTraceGroup = 64;

/* Waves per SIMD the tracers are built for: 4 for the FP32 and mixed ones (at most 128 VGPRs, which
 * they fit in; the mixed tile kernel measured 10% faster at 4 than at 3); the FP64 one would take 250
 * VGPRs for 2 waves, and is held to 168 for 3, which spills a little in its rarely taken paths and
 * measures 5% faster (4 spills too much). */
constexpr int wavesPerSimd(int precision) { return precision == ARTF_PRECISION_FP64 ? 3 : 4; }

/* The tile kernel: the same, but 5 for the FP32 tracers, which fit 96 VGPRs (the exact one with one value
 * spilled, stored as a ray starts and read as it ends; 6 waves would spill 12 and run 3% slower). More
 * waves to choose from raise the SIMDs' two-instructions-a-cycle issue (the kernel alone 2% faster); in a
 * frame, the post passes' waves (at most 32 VGPRs each) still fit in the 32 a SIMD then has left. */
constexpr int traceWavesPerSimd(int precision) {
    return precision == ARTF_PRECISION_FP32 || precision == ARTF_PRECISION_FP32_FAST ? 5 : wavesPerSimd(precision);
}

/* The GPU's clock (wall_clock64: artfHipWallClockHz ticks a second) stored to *at by the launch's first
 * lane of its first workgroup as it starts, with a system-scope store (at may be host memory): a stage
 * boundary read off the device's own clock, with no marker between kernels. NULL: nothing. */
__device__ inline void stampStart(unsigned long long* at, unsigned long long* also = nullptr) {
    if (at && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
        const unsigned long long now = (unsigned long long)wall_clock64();
        __hip_atomic_store(at, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (also) __hip_atomic_store(also, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); /* the same instant, two boundaries */
    }
}

/* FP32, the reference's arithmetic bit for bit (trace_exact.hip, contraction off). */
hipError_t launchTraceExact(const ArtfHipTraceArgs& a, hipStream_t stream);
/* FP32 with fused multiply-adds, FP64, and mixed (trace_fast.hip). */
hipError_t launchTraceFast(const ArtfHipTraceArgs& a, hipStream_t stream);
/* FP64 (trace_fp64.hip), which launchTraceFast, launchViewsFast and viewsOccupancyFast hand FP64 to. */
hipError_t launchTrace64(const ArtfHipTraceArgs& a, hipStream_t stream);
hipError_t launchViews64(const ArtfHipBatchArgs& a, hipStream_t stream);
int viewsOccupancy64();
/* The persistent tracer of a batch of views: FP32 (trace_exact.hip), the others (trace_fast.hip); and
 * the workgroups of it a compute unit holds. */
hipError_t launchViewsExact(const ArtfHipBatchArgs& a, hipStream_t stream);
hipError_t launchViewsFast(const ArtfHipBatchArgs& a, hipStream_t stream);
int viewsOccupancyExact();
int viewsOccupancyFast(int precision);
/* Clears a batch's workspace and stamps its start (post.hip). */
hipError_t launchBatchReset(const ArtfHipBatchArgs& a, hipStream_t stream);
hipError_t launchQueueReset(const ArtfHipBatchArgs& a, hipStream_t stream);
/* The disc map's painted gas, dust and sky (disc.hip, contraction off). */
hipError_t launchDiscPaint(float* map, hipStream_t stream);
/* The disc fluid's passes (fluid.hip, contraction off). */
hipError_t launchFluidSeed(const ArtfHipFluid& f, hipStream_t stream);
hipError_t launchFluidStir(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, int count, hipStream_t stream);
hipError_t launchFluidStep(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, hipStream_t stream);
hipError_t launchFluidField(const ArtfHipFluid& f, hipStream_t stream);
/* The analytic-disc adjoint (adjoint.hip, contraction off). */
hipError_t launchAdjoint(const ArtfHipAdjointArgs& a, hipStream_t stream);
int adjointBlocks(const ArtfHoleFrame& f);
/* Resolve, bloom and tone map (post.hip, contraction off). */
hipError_t launchResolve(const ArtfHoleFrame& f, float* const rays[3], float* const hdr[3], hipStream_t stream);
hipError_t launchBloom(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], hipStream_t stream);
hipError_t launchToneMap(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], uint32_t* bgra, hipStream_t stream);
/* A whole frame: trace, resolve with level 0, the rest of the pyramid, tone map (post.hip). */
hipError_t launchFrame(const ArtfHipFrameArgs& a, hipStream_t stream);

} // namespace artf
