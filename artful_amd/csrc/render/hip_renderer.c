/* artful-hip — the HIP backend: a renderer that draws a scene's black hole on an AMD Instinct GPU.
 *
 * The drop-in beside Artful 0.4.0's software and Vulkan backends (src/software/software.c:762-789,
 * src/vulkan/vulkan.c:1691-1783, 2566-2584): it fills the reference's vtable (rendererInternal.h:7-16)
 * and reads what they read of a scene (the map, uploaded when mapVersion changes; the fluid field, when
 * fluidVersion changes; the frame's uniforms from artfHoleFrame), through the calls of hip_scene.h, so
 * that this file compiles against this library's scene code or the reference's. Per frame it runs trace
 * (each pixel's rays summed in its lanes) -> bloom level 0 -> the rest of the bloom -> tone map
 * (artfHipDrawFrame), its stages timed by the GPU's clock, which the kernels stamp as they start. Frames
 * take turns with two sets of counters, HDR and pyramid, and each set has its own stream for its
 * traces; the blooms and tone maps go on a third. So a frame's trace starts as the last one's drains,
 * and a frame's post passes run beside the next frame's trace. A set is used again only after the
 * frame before last is done with it (its stream waits for that frame's tone map when it has not ended
 * yet), and the map and fluid are replaced only after the last frame's trace (the stream that uploads
 * waits for it). It is plain C over include/artful_hip.h.
 *
 * A draw only queues its frame (its step counts summed into pinned memory by its last kernel); the
 * frame is waited for by what needs it: reading pixels or HDR back, the stats, the next draw's host
 * work never. So frames follow one another on the device without the host in between, as the Vulkan
 * backend's frames in flight do (vulkan.c vkDraw waits on a fence some frames old). A device fault in
 * a queued frame is reported by the call that waits for it.
 *
 * Device memory per renderer: the map (11.5 MB) and fluid (1 MB) once; per size, two sets of HDR planes
 * (12 B/px) and bloom pyramids, and BGRA. */
#include "render/rendererInternal.h"

#include <stdio.h>

#include "artful/hip_renderer.h"
#include "artful_hip.h"
#include "hip_scene.h"

int artfHipCallsLeft = -1;

typedef struct HipRenderer {
    ArtfRenderer base;
    int device;
    void* traces[2]; /* per set: its frames' traces (and uploads) */
    void* post;      /* the blooms and tone maps */
    void* handoff[2], *done[2]; /* per set: its trace ended; its post passes ended */
    bool busy[2];   /* per set: a frame's post passes may still use it */
    int set;        /* the set of the last frame */
    double clockHz; /* the GPU clock's ticks a second */
    char description[256];
    int width, height;
    ArtfHoleTracing tracing;
    bool lost;
    /* scene data on the device, and which version of which scene it is */
    float* map;
    uint64_t mapScene, mapVersion;
    float* fluid;
    uint64_t fluidScene, fluidVersion;
    /* per-size targets */
    float* hdr[2][3];
    size_t hdrPixels;
    float* bloom[2][3];
    size_t bloomPixels;
    uint32_t* bgra;
    size_t bgraPixels;
    unsigned long long* counters[2]; /* device, per set: 2 * ARTF_HIP_COUNTER_SLOTS, zero between frames */
    unsigned long long* report;   /* pinned: the last frame's counts and clock stamps (ArtfHipFrameArgs), then
                                   * the first queued frame's trace start (ArtfHipTraceArgs.spanStamp) */
    /* the last frame */
    bool drawn, hole;
    bool queued;     /* its stats still to be read off the device */
    uint32_t frames; /* frames queued since the stats were last read */
    size_t frameRays;
    uint32_t background;
    ArtfHoleFrame frame;
    ArtfHoleStats holeStats;
    double cpuMs;
    /* the overlay, copied */
    uint32_t* overlay;
    int overlayWidth, overlayHeight;
} HipRenderer;

/* A device call's status: false (and the renderer lost) on failure, or when the test hook says so. */
static bool ok(HipRenderer* h, int status, const char* what) {
    if (status == 0 && artfHipCallsLeft != 0) {
        if (artfHipCallsLeft > 0) --artfHipCallsLeft;
        return true;
    }
    if (status) artfLog("artful: HIP %s failed: %s", what, artfHipErrorText(status));
    else artfLog("artful: HIP %s: the device is lost (test hook)", what);
    h->lost = true;
    return false;
}

static void freePlanes(float** p) {
    for (int c = 0; c < 3; ++c) {
        artfHipFree(p[c]);
        p[c] = NULL;
    }
}

static bool allocPlanes(HipRenderer* h, float** p, size_t floats) {
    for (int c = 0; c < 3; ++c) {
        p[c] = artfHipMalloc(floats * sizeof(float));
        if (!p[c]) {
            artfLog("artful: HIP: out of device memory for %zu floats", floats);
            freePlanes(p);
            h->lost = true;
            return false;
        }
    }
    return true;
}

static const char* hipDescription(const ArtfRenderer* r) { return ((const HipRenderer*)r)->description; }

static void hipResize(ArtfRenderer* r, int width, int height) {
    HipRenderer* h = (HipRenderer*)r;
    h->width = artfMaxi(width, 1);
    h->height = artfMaxi(height, 1);
    h->drawn = false;
}

static void hipSetOverlay(ArtfRenderer* r, const ArtfOverlay* o) {
    HipRenderer* h = (HipRenderer*)r;
    if (!o || !o->pixels) {
        artfFree(h->overlay);
        h->overlay = NULL;
        return;
    }
    const int stride = o->stride ? o->stride : o->width;
    if (!h->overlay || h->overlayWidth != o->width || h->overlayHeight != o->height) {
        artfFree(h->overlay);
        h->overlay = artfMalloc((size_t)o->width * o->height * 4);
        h->overlayWidth = o->width;
        h->overlayHeight = o->height;
    }
    for (int y = 0; y < o->height; ++y) memcpy(h->overlay + (size_t)y * o->width, o->pixels + (size_t)y * stride, (size_t)o->width * 4);
}

static bool settle(HipRenderer* h);

static bool ensure(HipRenderer* h, uint32_t bloomPixels) {
    const size_t pixels = (size_t)h->width * h->height;
    /* a queued frame may still use what is about to be replaced */
    const bool replace = pixels != h->hdrPixels || bloomPixels > h->bloomPixels;
    if (replace && h->queued && !settle(h)) return false;
    if (pixels != h->hdrPixels) {
        freePlanes(h->hdr[0]), freePlanes(h->hdr[1]);
        artfHipFree(h->bgra);
        h->bgra = NULL;
        h->hdrPixels = h->bgraPixels = 0;
        if (!allocPlanes(h, h->hdr[0], pixels) || !allocPlanes(h, h->hdr[1], pixels)) return false;
        h->bgra = artfHipMalloc(pixels * 4);
        if (!h->bgra) return ok(h, 2, "allocate");
        h->hdrPixels = h->bgraPixels = pixels;
    }
    if (bloomPixels > h->bloomPixels) {
        freePlanes(h->bloom[0]), freePlanes(h->bloom[1]);
        h->bloomPixels = 0;
        if (!allocPlanes(h, h->bloom[0], bloomPixels) || !allocPlanes(h, h->bloom[1], bloomPixels)) return false;
        h->bloomPixels = bloomPixels;
    }
    for (int k = 0; k < 2; ++k) {
        if (h->counters[k]) continue;
        const size_t bytes = 2 * ARTF_HIP_COUNTER_SLOTS * sizeof(unsigned long long);
        h->counters[k] = artfHipMalloc(bytes);
        if (!h->counters[k]) return ok(h, 2, "allocate");
        /* zero once: each frame's bloom level 0 leaves them zero for the set's next frame */
        if (!ok(h, artfHipMemset(h->counters[k], 0, bytes, h->traces[k]), "clear") || !ok(h, artfHipStreamSync(h->traces[k]), "clear")) return false;
    }
    return true;
}

/* Before the map or fluid the traces read is replaced on set's stream: that stream waits for the other
 * set's last trace (its own are in order already). */
static bool uploadAfterTraces(HipRenderer* h, int set) { return ok(h, artfHipStreamWait(h->traces[set], h->handoff[set ^ 1]), "upload order"); }

/* The map and the fluid on the device, as the scene now has them, uploaded on set's stream. */
static bool syncScene(HipRenderer* h, const ArtfHipSceneView* v, int set) {
    void* stream = h->traces[set];
    if (!h->map) {
        h->map = artfHipMalloc((size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float));
        h->fluid = artfHipMalloc((size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT * sizeof(float));
        if (!h->map || !h->fluid) return ok(h, 2, "allocate");
    }
    if (v->deviceMap && v->discDevice == h->device) {
        h->mapScene = 0; /* the scene's map is on this device: read in place */
    } else if (h->mapScene != v->id || h->mapVersion != v->mapVersion) {
        const size_t bytes = (size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float);
        if (!uploadAfterTraces(h, set)) return false;
        int status;
        if (v->deviceMap) { /* on another device: through the host */
            float* staged = artfMalloc(bytes);
            status = artfHipDownload(staged, v->deviceMap, bytes, NULL);
            if (!status) status = artfHipUpload(h->map, staged, bytes, stream);
            if (!status) status = artfHipStreamSync(stream);
            artfFree(staged);
        } else {
            status = artfHipUpload(h->map, v->map, bytes, stream);
        }
        if (!ok(h, status, "map upload")) return false;
        h->mapScene = v->id;
        h->mapVersion = v->mapVersion;
    }
    if (v->deviceFluid && v->discDevice == h->device) {
        /* the scene steps its fluid on this device: read it where it is (the draw orders its trace after
         * the current field's write and marks its read) */
        h->fluidScene = 0;
        return true;
    }
    if (h->fluidScene != v->id || h->fluidVersion != v->fluidVersion) {
        const size_t bytes = (size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT * sizeof(float);
        if (!uploadAfterTraces(h, set)) return false;
        int status;
        if (v->deviceFluid) { /* on another device: through the host */
            float* field = artfMalloc(bytes);
            status = v->deviceFluidRead(v->deviceFluid, field) ? artfHipUpload(h->fluid, field, bytes, stream) : 1;
            if (!status) status = artfHipStreamSync(stream);
            artfFree(field);
        } else {
            /* from the scene's pageable field: the copy has read it when this returns (the runtime
             * stages pageable sources), so the scene's next fluid step may rewrite it at once */
            status = artfHipUpload(h->fluid, v->fluid, bytes, stream);
        }
        if (!ok(h, status, "fluid upload")) return false;
        h->fluidScene = v->id;
        h->fluidVersion = v->fluidVersion;
    }
    return true;
}

/* The frame's uniforms with at most `cap` bloom levels (0: the reference's rule). The levels are
 * halvings from the first, so fewer of them is the table cut short. */
static void frameFor(const ArtfScene* scene, int width, int height, int cap, ArtfHoleFrame* f, uint32_t* bloomPixels) {
    artfHipSceneFrame(scene, width, height, f, bloomPixels);
    if (cap > 0 && (uint32_t)cap < f->size[3]) {
        const uint32_t* last = f->bloom[cap - 1];
        *bloomPixels = last[0] + last[1] * last[2];
        for (int k = cap; k < ARTF_BLOOM_LEVELS; ++k) f->bloom[k][0] = f->bloom[k][1] = f->bloom[k][2] = f->bloom[k][3] = 0;
        f->size[3] = (uint32_t)cap;
    }
}

static bool hipDraw(ArtfRenderer* r, ArtfScene* scene) {
    HipRenderer* h = (HipRenderer*)r;
    if (h->lost) return false;
    const double t0 = artfNowMs();
    ArtfHipSceneView view;
    artfHipSceneView(scene, &view);
    h->drawn = false;
    h->hole = view.hole;
    memset(&h->holeStats, 0, sizeof h->holeStats);
    if (!h->hole) {
        /* the scene's background, as the reference's frame without a hole (software.c:597-598) */
        h->background = view.background;
        h->drawn = true;
        h->cpuMs = artfNowMs() - t0;
        return true;
    }
    uint32_t bloomPixels = 0;
    frameFor(scene, h->width, h->height, h->tracing.bloomLevels, &h->frame, &bloomPixels);
    const int side = (int)h->frame.size[2];
    const size_t rays = (size_t)h->width * h->height * side * side;
    if (h->height > 65535 || rays >= ((size_t)1 << 31)) {
        /* beyond a launch grid (the renderer stays usable: a smaller size draws) */
        artfLog("artful: HIP: a %dx%d frame at %d rays a pixel is beyond the kernels' grids (at most 65535 rows, under 2^31 rays)",
                h->width, h->height, side * side);
        return false;
    }
    if (!ensure(h, bloomPixels)) return false;
    /* the other set than the last frame's, once the frame before last is done with it */
    const int set = h->set ^ 1;
    if (!syncScene(h, &view, set)) return false;

    ArtfHipFrameArgs fa;
    memset(&fa, 0, sizeof fa);
    ArtfHipTraceArgs* a = &fa.trace;
    a->frame = h->frame;
    if (h->tracing.precision == ARTF_PRECISION_MIXED || h->tracing.precision == ARTF_PRECISION_FP64) artfHipSceneFrame64(scene, h->width, h->height, &a->frame64);
    a->map = view.deviceMap && view.discDevice == h->device ? view.deviceMap : h->map;
    /* a fluid stepped on this device: its field read in place, once written (its steps are queued on its
     * own stream), and not written again in that buffer before this trace has read it */
    void* deviceFluid = view.deviceFluid && view.discDevice == h->device ? view.deviceFluid : NULL;
    a->fluid = deviceFluid ? view.deviceFluidField(deviceFluid) : h->fluid;
    if (deviceFluid && !ok(h, view.deviceFluidReadOn(deviceFluid, h->traces[set]) ? 0 : 1, "fluid order")) return false;
    if (h->busy[set]) {
        const int done = artfHipEventDone(h->done[set]);
        if (done < 0) return ok(h, 1, "frame order");
        if (!done && !ok(h, artfHipStreamWait(h->traces[set], h->done[set]), "frame order")) return false;
    }
    for (int c = 0; c < 3; ++c) a->out[c] = fa.hdr[c] = h->hdr[set][c], fa.bloom[c] = h->bloom[set][c];
    a->counters = h->tracing.uncounted ? NULL : h->counters[set];
    a->precision = h->tracing.precision;
    a->maxSteps = h->tracing.maxSteps;
    a->escapeRadius = h->tracing.escapeRadius;
    a->disc = h->tracing.disc;
    a->emissivity = h->tracing.emissivity;
    a->jitter = h->tracing.jitter;
    /* the first frame queued since the stats were read marks the start of their span */
    a->spanStamp = h->frames == 0 ? h->report + ARTF_HIP_REPORT_WORDS : NULL;
    fa.bgra = h->bgra;
    fa.report = h->report;
    fa.postStream = h->post;
    fa.handoff = h->handoff[set];
    fa.done = h->done[set];
    if (!ok(h, artfHipDrawFrame(&fa, h->traces[set]), "frame")) return false;
    if (deviceFluid && !ok(h, view.deviceFluidReadDone(deviceFluid, h->traces[set]) ? 0 : 1, "fluid order")) return false;
    h->busy[set] = true;
    h->set = set;
    h->frameRays = rays;
    h->queued = true;
    ++h->frames;
    h->drawn = true;
    h->cpuMs = artfNowMs() - t0;
    return true;
}

/* The queued frames done (both streams): the last one's stage times and step counts from its report. */
static bool settle(HipRenderer* h) {
    if (!h->queued) return !h->lost;
    h->queued = false;
    for (int k = 0; k < 2; ++k)
        if (!ok(h, artfHipStreamSync(h->traces[k]), "frame")) return false;
    if (!ok(h, artfHipStreamSync(h->post), "frame")) return false;
    h->busy[0] = h->busy[1] = false;
    ArtfHoleStats* st = &h->holeStats;
    /* each stage from its kernel's start to the next one's (the tone map: to its last workgroup's end);
     * the trace sums each pixel's rays itself, so there is no resolve to report */
    const unsigned long long* r = h->report;
    const double ms = 1e3 / h->clockHz;
    st->traceMs = (double)(r[3] - r[2]) * ms;
    st->resolveMs = 0;
    st->bloomMs = (double)(r[5] - r[3]) * ms;
    st->toneMs = (double)(r[6] - r[5]) * ms;
    st->uploadMs = 0;
    st->rays = h->frameRays;
    st->steps = r[0];
    st->laneSlots = r[1];
    st->bloomLevels = (int)h->frame.size[3];
    /* the frames queued since the last wait, from the first one's trace start to the last one's end */
    st->frames = h->frames;
    st->spanMs = (double)(r[6] - r[ARTF_HIP_REPORT_WORDS]) * ms;
    h->frames = 0;
    return true;
}

/* The overlay over the frame as the reference puts it there (kernels.inc:1264-1272, 1741-1756): each
 * channel, alpha too, d = s + d (255 - sa) / 255, the division rounded ((t + (t >> 8)) >> 8 of
 * t = d (255 - sa) + 128); a pixel that is 0 as a whole leaves the frame as it is (and so would the
 * formula), one of alpha 0 and some colour adds its colour (a premultiplied glow). */
static void compose(const HipRenderer* h, uint32_t* dst) {
    if (!h->overlay) return;
    const int w = artfMini(h->overlayWidth, h->width), hh = artfMini(h->overlayHeight, h->height);
    for (int y = 0; y < hh; ++y)
        for (int x = 0; x < w; ++x) {
            const uint32_t s = h->overlay[(size_t)y * h->overlayWidth + x];
            if (!s) continue;
            uint32_t* p = dst + (size_t)y * h->width + x;
            const uint32_t rest = 255 - (s >> 24);
            uint32_t blended = 0;
            for (int c = 0; c < 32; c += 8) {
                const uint32_t t = ((*p >> c) & 255) * rest + 128;
                blended |= ((t + (t >> 8)) >> 8) << c;
            }
            *p = s + blended;
        }
}

static bool hipReadPixels(ArtfRenderer* r, uint32_t* dst) {
    HipRenderer* h = (HipRenderer*)r;
    if (h->lost || !h->drawn || !settle(h)) return false;
    const size_t pixels = (size_t)h->width * h->height;
    if (!h->hole) {
        for (size_t i = 0; i < pixels; ++i) dst[i] = h->background;
    } else if (!ok(h, artfHipDownload(dst, h->bgra, pixels * 4, h->post), "read back")) {
        return false;
    }
    compose(h, dst);
    return true;
}

static uint32_t* hipPixels(ArtfRenderer* r) {
    (void)r;
    return NULL;
}

static void hipStats(const ArtfRenderer* r, ArtfFrameStats* out) {
    HipRenderer* h = (HipRenderer*)r; /* (settling the queued frame fills in what is read) */
    settle(h);
    const ArtfHoleStats* st = &h->holeStats;
    out->cpuMs = h->cpuMs;
    out->lensMs = st->traceMs + st->resolveMs + st->bloomMs;
    out->rasterMs = st->toneMs;
    out->gpuMs = out->lensMs + out->rasterMs;
}

static void hipDestroy(ArtfRenderer* r) {
    HipRenderer* h = (HipRenderer*)r;
    for (int k = 0; k < 2; ++k)
        if (h->traces[k]) artfHipStreamSync(h->traces[k]);
    if (h->post) artfHipStreamSync(h->post);
    for (int k = 0; k < 2; ++k) {
        freePlanes(h->hdr[k]), freePlanes(h->bloom[k]);
        artfHipFree(h->counters[k]);
        artfHipEventDestroy(h->handoff[k]), artfHipEventDestroy(h->done[k]);
    }
    artfHipFree(h->bgra);
    artfHipFree(h->map);
    artfHipFree(h->fluid);
    artfHipStreamDestroy(h->traces[0]), artfHipStreamDestroy(h->traces[1]);
    artfHipStreamDestroy(h->post);
    artfFree(h->overlay);
    artfHipHostFree(h->report);
    artfFree(h);
}

static const ArtfRendererVtbl hipTable = {hipDescription, hipResize, hipSetOverlay, hipDraw, hipReadPixels, hipPixels, hipStats, hipDestroy};

/* ---- beyond the reference (artful/hip_renderer.h): any renderer may be given; another backend's is
 * left alone ---- */

static HipRenderer* asHip(const ArtfRenderer* r, const char* call) {
    if (r && r->vt == &hipTable) return (HipRenderer*)r;
    artfLog("artful: %s: not a HIP renderer", call);
    return NULL;
}

bool artfRendererReadHdr(ArtfRenderer* r, float* red, float* green, float* blue) {
    HipRenderer* h = asHip(r, "artfRendererReadHdr");
    if (!h || h->lost || !h->drawn || !h->hole || !settle(h)) return false;
    const size_t bytes = (size_t)h->width * h->height * sizeof(float);
    float* dst[3] = {red, green, blue};
    for (int c = 0; c < 3; ++c)
        if (!ok(h, artfHipDownload(dst[c], h->hdr[h->set][c], bytes, h->post), "read back")) return false;
    return true;
}

ArtfHoleTracing artfHoleTracingDefaults(void) {
    ArtfHoleTracing t;
    memset(&t, 0, sizeof t);
    t.precision = ARTF_PRECISION_FP32;
    t.disc = ARTF_DISC_PAINTED;
    t.emissivity = 3.f;
    return t;
}

void artfRendererSetHoleTracing(ArtfRenderer* r, const ArtfHoleTracing* t) {
    HipRenderer* h = asHip(r, "artfRendererSetHoleTracing");
    if (!h) return;
    h->tracing = t ? *t : artfHoleTracingDefaults();
    h->drawn = false;
}

bool artfRendererHoleStats(const ArtfRenderer* r, ArtfHoleStats* out) {
    memset(out, 0, sizeof *out);
    HipRenderer* h = asHip(r, "artfRendererHoleStats");
    if (!h) return false;
    const bool settled = settle(h);
    *out = h->holeStats;
    return settled && h->drawn && h->hole;
}

ArtfRenderer* artfHipRendererNew(const ArtfRendererOptions* o, int device) {
    const int count = artfHipDeviceCount();
    if (count <= 0 || device < 0 || device >= count) {
        artfLog("artful: no HIP device %d here (%d found)", device, count);
        return NULL;
    }
    if (artfHipSetDevice(device)) {
        artfLog("artful: HIP device %d cannot be used", device);
        return NULL;
    }
    HipRenderer* h = ARTF_NEW(HipRenderer);
    h->base.vt = &hipTable;
    h->base.backend = (ArtfBackend)ARTF_HIP_BACKEND_VALUE;
    h->device = device;
    h->traces[0] = artfHipStreamCreate();
    h->traces[1] = artfHipStreamCreate();
    h->post = artfHipStreamCreate();
    bool ready = h->traces[0] && h->traces[1] && h->post;
    for (int k = 0; k < 2; ++k) {
        h->handoff[k] = artfHipEventCreateOrder();
        h->done[k] = artfHipEventCreateOrder();
        ready = ready && h->handoff[k] && h->done[k];
    }
    h->clockHz = artfHipWallClockHz(device);
    if (!ready || !(h->clockHz > 0)) {
        artfLog("artful: HIP device %d: no streams, events or clock", device);
        hipDestroy(&h->base);
        return NULL;
    }
    char name[160] = "";
    int units = 0;
    artfHipDeviceInfo(device, name, sizeof name, &units);
    snprintf(h->description, sizeof h->description, "HIP, %s, %d compute units", name, units);
    h->tracing = artfHoleTracingDefaults();
    /* the frame's report, then the word the first of a run of frames stamps its start into */
    const size_t words = ARTF_HIP_REPORT_WORDS + 1;
    h->report = artfHipHostMalloc(words * sizeof *h->report);
    if (!h->report) {
        artfLog("artful: HIP device %d: no pinned host memory", device);
        hipDestroy(&h->base);
        return NULL;
    }
    memset(h->report, 0, words * sizeof *h->report);
    hipResize(&h->base, o->width, o->height);
    return &h->base;
}

ArtfRenderer* artfHipNew(const ArtfRendererOptions* o) { return artfHipRendererNew(o, 0); }
