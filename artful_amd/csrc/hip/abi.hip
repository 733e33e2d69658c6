/* artful-hip — the C ABI over the kernels (include/artful_hip.h). Every entry checks what it is given
 * on the host before anything is launched, and returns a HIP status; nothing here allocates except
 * artfHipMalloc. */
#include <stdio.h>
#include <string.h>

#include "launch.h"

using namespace artf;

extern "C" {

int artfHipDeviceCount(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int artfHipSetDevice(int device) { return (int)hipSetDevice(device); }

int artfHipDeviceInfo(int device, char* name, size_t size, int* computeUnits) {
    hipDeviceProp_t p;
    const hipError_t e = hipGetDeviceProperties(&p, device);
    if (e != hipSuccess) return (int)e;
    /* (the marketing name can be empty, as under some drivers: then the architecture alone) */
    if (name && size) snprintf(name, size, p.name[0] ? "%s (%s)" : "%s%s", p.name, p.gcnArchName);
    if (computeUnits) *computeUnits = p.multiProcessorCount;
    return 0;
}

void* artfHipMalloc(size_t bytes) {
    void* p = NULL;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return NULL;
    return p;
}

void artfHipFree(void* p) {
    if (p) (void)hipFree(p);
}

int artfHipUpload(void* dst, const void* src, size_t bytes, void* stream) {
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
}

int artfHipDownload(void* dst, const void* src, size_t bytes, void* stream) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e != hipSuccess) return (int)e;
    return (int)hipStreamSynchronize((hipStream_t)stream);
}

int artfHipDownloadAsync(void* dst, const void* src, size_t bytes, void* stream) {
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
}

void* artfHipHostMalloc(size_t bytes) {
    void* p = NULL;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocCoherent) != hipSuccess) return NULL;
    return p;
}

void artfHipHostFree(void* p) {
    if (p) (void)hipHostFree(p);
}

int artfHipMemset(void* dst, int value, size_t bytes, void* stream) {
    return (int)hipMemsetAsync(dst, value, bytes, (hipStream_t)stream);
}

int artfHipCopy(void* dst, const void* src, size_t bytes, void* stream) {
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
}

void* artfHipStreamCreate(void) {
    hipStream_t s = NULL;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return NULL;
    return (void*)s;
}

void artfHipStreamDestroy(void* stream) {
    if (stream) (void)hipStreamDestroy((hipStream_t)stream);
}

int artfHipStreamSync(void* stream) { return (int)hipStreamSynchronize((hipStream_t)stream); }

int artfHipDeviceSync(void) { return (int)hipDeviceSynchronize(); }

void* artfHipEventCreate(void) {
    hipEvent_t e = NULL;
    if (hipEventCreate(&e) != hipSuccess) return NULL;
    return (void*)e;
}

void artfHipEventDestroy(void* event) {
    if (event) (void)hipEventDestroy((hipEvent_t)event);
}

int artfHipEventRecord(void* event, void* stream) { return (int)hipEventRecord((hipEvent_t)event, (hipStream_t)stream); }

void* artfHipEventCreateOrder(void) {
    hipEvent_t e = NULL;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventReleaseToDevice) != hipSuccess) return NULL;
    return (void*)e;
}

int artfHipStreamWait(void* stream, void* event) { return (int)hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0); }

int artfHipEventDone(void* event) {
    const hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) {
        (void)hipGetLastError(); /* not an error: leave nothing for the next launch's check */
        return 0;
    }
    return -1;
}

float artfHipEventMs(void* start, void* stop) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop) != hipSuccess) return -1.f;
    return ms;
}

const char* artfHipErrorText(int status) { return hipGetErrorString((hipError_t)status); }

/* The stream of a launch entry, with the thread's error state cleared: a launcher's status is
 * hipGetLastError() after its launches, which must not report a call that failed before (an event
 * never recorded, a refused copy) as this launch's. */
static hipStream_t fresh(void* stream) {
    (void)hipGetLastError();
    return (hipStream_t)stream;
}

/* A frame the kernels can take: at most 65535 rows (the resolve and tone map launch a row of
 * workgroups per row of pixels, and a grid is at most 65535 workgroups high), fewer than 2^31 rays. */
static bool frameFits(const ArtfHoleFrame* f) {
    return f && f->size[0] > 0 && f->size[1] > 0 && f->size[1] <= 65535 && f->size[2] >= 1 && f->size[2] <= 8 && f->size[3] >= 1 &&
           f->size[3] <= ARTF_BLOOM_LEVELS && (uint64_t)f->size[0] * f->size[1] * f->size[2] * f->size[2] < (1ull << 31);
}

static bool traceFits(const ArtfHipTraceArgs* a) {
    return a && frameFits(&a->frame) && a->map && a->fluid && a->out[0] && a->out[1] && a->out[2] && a->maxSteps >= 0 && a->maxSteps <= 1000000 &&
           a->rowBegin >= 0;
}

/* The bloom level table artfHoleFrame builds: each level half the last (rounded up), packed. */
static bool pyramidFits(const ArtfHoleFrame* f) {
    uint32_t w = f->size[0], h = f->size[1], at = 0;
    for (uint32_t k = 0; k < f->size[3]; ++k) {
        w = (w + 1) / 2, h = (h + 1) / 2;
        if (f->bloom[k][0] != at || f->bloom[k][1] != w || f->bloom[k][2] != h) return false;
        at += w * h;
    }
    return true;
}

int artfHipTrace(const ArtfHipTraceArgs* a, void* stream) {
    if (!traceFits(a)) return (int)hipErrorInvalidValue;
    if (a->precision == ARTF_PRECISION_FP32) return (int)launchTraceExact(*a, fresh(stream));
    return (int)launchTraceFast(*a, fresh(stream));
}

int artfHipAdjointBlocks(const ArtfHoleFrame* f) { return frameFits(f) ? adjointBlocks(*f) : 0; }

int artfHipAdjoint(const ArtfHipAdjointArgs* args, void* stream) {
    if (!args) return (int)hipErrorInvalidValue;
    const ArtfHipTraceArgs* a = &args->trace;
    if (!frameFits(&a->frame) || !a->map || !a->out[0] || !a->out[1] || !a->out[2] || !args->seed[0] || !args->seed[1] || !args->seed[2] ||
        !args->partials)
        return (int)hipErrorInvalidValue;
    if (a->disc != ARTF_DISC_ANALYTIC || (a->precision != ARTF_PRECISION_FP32 && a->precision != ARTF_PRECISION_FP64)) return (int)hipErrorInvalidValue;
    if (a->maxSteps < 0 || a->maxSteps > 2048 || a->frame.look[2] > 0.f || a->frame.boost[3] > 1.f || a->rowBegin != 0 || a->rowEnd != 0)
        return (int)hipErrorInvalidValue;
    return (int)launchAdjoint(*args, fresh(stream));
}

int artfHipResolve(const ArtfHoleFrame* f, float* const rays[3], float* const hdr[3], void* stream) {
    if (!frameFits(f) || !rays || !hdr) return (int)hipErrorInvalidValue;
    return (int)launchResolve(*f, rays, hdr, fresh(stream));
}

int artfHipBloom(const ArtfHoleFrame* f, float* const hdr[3], float* const bloom[3], void* stream) {
    if (!frameFits(f) || !hdr || !bloom || !pyramidFits(f)) return (int)hipErrorInvalidValue;
    return (int)launchBloom(*f, hdr, bloom, fresh(stream));
}

int artfHipToneMap(const ArtfHoleFrame* f, float* const hdr[3], float* const bloom[3], uint32_t* bgra, void* stream) {
    if (!frameFits(f) || !hdr || !bloom || !bgra) return (int)hipErrorInvalidValue;
    if (f->bloom[0][1] != (f->size[0] + 1) / 2 || f->bloom[0][2] != (f->size[1] + 1) / 2) return (int)hipErrorInvalidValue;
    return (int)launchToneMap(*f, hdr, bloom, bgra, fresh(stream));
}

int artfHipDrawFrame(const ArtfHipFrameArgs* a, void* stream) {
    if (!a || !traceFits(&a->trace) || !pyramidFits(&a->trace.frame) || !a->bgra || (a->report && !a->trace.counters)) return (int)hipErrorInvalidValue;
    const ArtfHoleFrame* f = &a->trace.frame;
    if (a->trace.rowBegin != 0 || (a->trace.rowEnd != 0 && a->trace.rowEnd < (int)f->size[1])) return (int)hipErrorInvalidValue; /* whole frames */
    for (int c = 0; c < 3; ++c)
        if (!a->hdr[c] || !a->bloom[c] || a->trace.out[c] != a->hdr[c]) return (int)hipErrorInvalidValue;
    if (a->postStream && (!a->handoff || a->postStream == stream)) return (int)hipErrorInvalidValue;
    return (int)launchFrame(*a, fresh(stream));
}

static bool batchFits(const ArtfHipBatchArgs* a) {
    if (!a || !a->views || !a->firstTile || !a->map || !a->rings || !a->queue || a->viewCount < 1 || a->tileBegin >= a->tileEnd ||
        a->tileEnd > (1u << 25))
        return false;
    if (a->viewCount > 65536) return false;
    const ArtfHipSchedule& s = a->schedule;
    if (s.workgroups < 1 || s.workgroups > 65536 || s.chunkTiles < 1 || s.chunkTiles > 64 || s.refill < 1 || s.refill > 64 || s.stepsPerCheck < 1 ||
        s.stepsPerCheck > 16)
        return false;
    return a->precision >= ARTF_PRECISION_FP32 && a->precision <= ARTF_PRECISION_FP64;
}

int artfHipBatchReset(const ArtfHipBatchArgs* a, void* stream) {
    if (!batchFits(a)) return (int)hipErrorInvalidValue;
    return (int)launchBatchReset(*a, fresh(stream));
}

/* The views themselves live on the device, where the host cannot check them: the caller (the batch in
 * render/hole_batch.c) builds them from frames it has checked with frameFits and uploads them. */
int artfHipBatchResetQueues(const ArtfHipBatchArgs* a, void* stream) {
    if (!batchFits(a)) return (int)hipErrorInvalidValue;
    return (int)launchQueueReset(*a, fresh(stream));
}

int artfHipTraceViews(const ArtfHipBatchArgs* a, void* stream) {
    if (!batchFits(a)) return (int)hipErrorInvalidValue;
    if (a->precision == ARTF_PRECISION_FP32) return (int)launchViewsExact(*a, fresh(stream));
    return (int)launchViewsFast(*a, fresh(stream));
}

int artfHipViewsOccupancy(int precision) {
    if (precision == ARTF_PRECISION_FP32) return viewsOccupancyExact();
    return viewsOccupancyFast(precision);
}

double artfHipWallClockHz(int device) {
    int kHz = 0;
    if (hipDeviceGetAttribute(&kHz, hipDeviceAttributeWallClockRate, device) != hipSuccess) return 0.0;
    return (double)kHz * 1e3;
}

int artfHipGraphBegin(void* stream) { return (int)hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal); }

int artfHipGraphEnd(void* stream, void** graph) {
    if (!graph) return (int)hipErrorInvalidValue;
    *graph = NULL;
    hipGraph_t g = NULL;
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
    if (e != hipSuccess) return (int)e;
    hipGraphExec_t exec = NULL;
    e = hipGraphInstantiate(&exec, g, NULL, NULL, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return (int)e;
    *graph = (void*)exec;
    return 0;
}

int artfHipGraphLaunch(void* graph, void* stream) {
    if (!graph) return (int)hipErrorInvalidValue;
    return (int)hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream);
}

void artfHipGraphFree(void* graph) {
    if (graph) (void)hipGraphExecDestroy((hipGraphExec_t)graph);
}

static bool fluidFits(const ArtfHipFluid* f) {
    return f && f->density && f->radial && f->around && f->scratch[0] && f->scratch[1] && f->scratch[2] && f->field && f->rings;
}

int artfHipFluidSeed(const ArtfHipFluid* f, void* stream) {
    if (!fluidFits(f)) return (int)hipErrorInvalidValue;
    return (int)launchFluidSeed(*f, fresh(stream));
}

int artfHipFluidStir(const ArtfHipFluid* f, const ArtfFluidEddy* eddies, int count, void* stream) {
    if (!fluidFits(f) || !eddies || count < 0) return (int)hipErrorInvalidValue;
    if (count == 0) return 0;
    return (int)launchFluidStir(*f, eddies, count, fresh(stream));
}

int artfHipFluidStep(const ArtfHipFluid* f, const ArtfFluidEddy* eddies, void* stream) {
    if (!fluidFits(f) || !eddies) return (int)hipErrorInvalidValue;
    return (int)launchFluidStep(*f, eddies, fresh(stream));
}

int artfHipFluidField(const ArtfHipFluid* f, void* stream) {
    if (!fluidFits(f)) return (int)hipErrorInvalidValue;
    return (int)launchFluidField(*f, fresh(stream));
}

int artfHipDiscPaint(float* map, void* stream) {
    if (!map) return (int)hipErrorInvalidValue;
    return (int)launchDiscPaint(map, fresh(stream));
}

} // extern "C"
