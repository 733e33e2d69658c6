/* artful-hip — the C ABI over the kernels (include/artful_hip.h). Every entry checks what it is given
 * on the host before anything is launched, and returns a HIP status; nothing here allocates except
 * artfHipMalloc (and artfHipHostMalloc, the streams, events and graphs it is asked for). */
#include <pthread.h>
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
    if (!a || !traceFits(&a->trace) || !pyramidFits(&a->trace.frame) || !a->bgra) return (int)hipErrorInvalidValue;
    const ArtfHoleFrame* f = &a->trace.frame;
    if (a->trace.rowBegin != 0 || (a->trace.rowEnd != 0 && a->trace.rowEnd < (int)f->size[1])) return (int)hipErrorInvalidValue; /* whole frames */
    for (int c = 0; c < 3; ++c)
        if (!a->hdr[c] || !a->bloom[c] || a->trace.out[c] != a->hdr[c]) return (int)hipErrorInvalidValue;
    if (a->postStream && (!a->handoff || a->postStream == stream)) return (int)hipErrorInvalidValue;
    return (int)launchFrame(*a, fresh(stream));
}

} // extern "C"

/* ---- artfHipDrawFrameStatic: artfHipDrawFrame with its scratch in a static device array ----
 *
 * A frame's only scratch is the trace's step counters (2 x ARTF_HIP_COUNTER_SLOTS words, 16 KiB): the
 * trace adds to them, bloom level 0's first workgroup sums them into the report and zeroes them. Here
 * they live in `scratch`, one copy per GPU for as long as the process runs, so that a caller that may not
 * allocate (under graph capture, in a serving loop) can still draw frames. Its size does not depend on
 * the frame: a slot is a workgroup's index modulo ARTF_HIP_COUNTER_SLOTS. In host code `scratch` names a
 * host-side stand-in, never device memory: the device address comes from hipGetSymbolAddress, once per
 * device with that device current, and only that address is handed to kernels. */
static __device__ __align__(256) unsigned long long scratch[2 * ARTF_HIP_COUNTER_SLOTS];

static constexpr int kScratchDevices = 64;
static unsigned long long* scratchAt[kScratchDevices];
static bool scratchKnown[kScratchDevices];
static pthread_mutex_t scratchLock = PTHREAD_MUTEX_INITIALIZER;

/* The current device's copy of `scratch` (NULL, with *status set, if it cannot be had). */
static unsigned long long* scratchOfCurrent(hipError_t* status) {
    int device = 0;
    if ((*status = hipGetDevice(&device)) != hipSuccess) return nullptr;
    if (device < 0 || device >= kScratchDevices) {
        *status = hipErrorInvalidDevice;
        return nullptr;
    }
    pthread_mutex_lock(&scratchLock);
    if (!scratchKnown[device]) {
        void* at = nullptr;
        *status = hipGetSymbolAddress(&at, HIP_SYMBOL(scratch));
        if (*status == hipSuccess && at) {
            scratchAt[device] = (unsigned long long*)at;
            scratchKnown[device] = true;
        } else if (*status == hipSuccess) {
            *status = hipErrorInvalidSymbol;
        }
    }
    unsigned long long* at = scratchKnown[device] ? scratchAt[device] : nullptr;
    pthread_mutex_unlock(&scratchLock);
    return at;
}

/* Zeroes the counters by a kernel (no memset node: the call may be captured into a graph). */
__global__ __launch_bounds__(256) void clearScratchKernel(unsigned long long* at) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < 2 * ARTF_HIP_COUNTER_SLOTS) at[i] = 0ull;
}

extern "C" {

int artfHipDrawFrameStatic(const ArtfHipFrameArgs* a, void* stream) {
    if (!a) return (int)hipErrorInvalidValue;
    ArtfHipFrameArgs b = *a;
    b.trace.counters = nullptr; /* whatever was given: the static array stands in for them */
    if (!traceFits(&b.trace) || !pyramidFits(&b.trace.frame) || !b.bgra) return (int)hipErrorInvalidValue;
    const ArtfHoleFrame* f = &b.trace.frame;
    if (b.trace.rowBegin != 0 || (b.trace.rowEnd != 0 && b.trace.rowEnd < (int)f->size[1])) return (int)hipErrorInvalidValue;
    for (int c = 0; c < 3; ++c)
        if (!b.hdr[c] || !b.bloom[c] || b.trace.out[c] != b.hdr[c]) return (int)hipErrorInvalidValue;
    if (b.postStream) return (int)hipErrorInvalidValue; /* one stream: a frame's counters are read before the next frame's trace adds */
    if (b.report) {
        hipError_t status;
        unsigned long long* at = scratchOfCurrent(&status);
        if (!at) return (int)status;
        hipStream_t s = fresh(stream);
        hipLaunchKernelGGL(clearScratchKernel, dim3((2 * ARTF_HIP_COUNTER_SLOTS + 255) / 256), dim3(256), 0, s, at);
        if ((status = hipGetLastError()) != hipSuccess) return (int)status;
        b.trace.counters = at;
        return (int)launchFrame(b, s);
    }
    return (int)launchFrame(b, fresh(stream));
}

static bool batchFits(const ArtfHipBatchArgs* a) {
    if (!a || !a->views || !a->firstTile || !a->map || !a->rings || !a->queue || a->viewCount < 1 || a->tileBegin >= a->tileEnd ||
        a->tileEnd > (1u << 25))
        return false;
    if (a->viewCount > 65536) return false;
    const ArtfHipSchedule& s = a->schedule;
    if (s.workgroups != ARTF_HIP_WORKGROUP_A_TILE && (s.workgroups < 1 || s.workgroups > 65536)) return false;
    if (s.chunkTiles < 1 || s.chunkTiles > 64 || s.refill < 1 || s.refill > 64 || s.stepsPerCheck < 1 || s.stepsPerCheck > 16) return false;
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
