/* artful-hip — drawing a scene's black hole on an MI355X (gfx950).
 *
 * The renderer API of Artful 0.4.0 (include/artful/renderer.h:24-133): the same names, enum values,
 * struct layouts and contracts, byte for byte (sizes and offsets are checked against the reference's
 * own build), with one more backend, ARTF_BACKEND_HIP, appended after the reference's values so they
 * keep their numbers. artfRendererDraw returns false when the device is lost (free the renderer, keep
 * the scene); artfRendererReadPixels writes width*height premultiplied BGRA8; artfRendererPixels is
 * NULL for a GPU.
 *
 * One deviation: the reference's AUTO falls back to its software backend where no GPU runs
 * (renderer.h:79, renderer.c:24), so it never gives NULL. This library has the HIP backend only, so
 * here AUTO gives NULL (logged) where no HIP device runs. Built into the reference's tree (its src/hip)
 * the backend leaves the reference's dispatch, and that fallback, as they are.
 *
 * What the HIP backend adds (HDR readback, precision and ray law, per-stage GPU times, the choice of
 * a device) is declared in artful/hip_renderer.h, included at the end. */
#ifndef ARTFUL_RENDERER_H
#define ARTFUL_RENDERER_H

#include "artful/scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ArtfBackend {
    ARTF_BACKEND_AUTO,     /* HIP here (the reference: Vulkan where it works, then software) */
    ARTF_BACKEND_VULKAN,   /* not in this library */
    ARTF_BACKEND_SOFTWARE, /* not in this library */
    ARTF_BACKEND_METAL,    /* not in this library */
    ARTF_BACKEND_HIP       /* hand-written HIP kernels on an AMD Instinct GPU */
} ArtfBackend;

const char* artfBackendName(ArtfBackend b);

typedef enum ArtfSurfaceKind {
    ARTF_SURFACE_NONE,    /* frames stay in memory: artfRendererReadPixels (the only kind the HIP backend draws to) */
    ARTF_SURFACE_WIN32,   /* a: the HWND */
    ARTF_SURFACE_XLIB,    /* a: the Display*, b: the Window */
    ARTF_SURFACE_WAYLAND, /* a: the wl_display*, b: the wl_surface* */
    ARTF_SURFACE_ANDROID, /* a: the ANativeWindow* */
    ARTF_SURFACE_METAL    /* a: the CAMetalLayer* */
} ArtfSurfaceKind;

typedef struct ArtfSurface {
    ArtfSurfaceKind kind;
    void* a;
    void* b;
} ArtfSurface;

typedef enum ArtfSimd { ARTF_SIMD_AUTO, ARTF_SIMD_PORTABLE, ARTF_SIMD_128, ARTF_SIMD_256 } ArtfSimd;

typedef struct ArtfRendererOptions {
    ArtfBackend backend; /* AUTO: HIP */
    ArtfSurface surface; /* the HIP backend draws into memory whatever is given */
    int width, height;
    bool noVsync;
    int threads;         /* the reference's software renderer's threads; the HIP backend draws on the GPU */
    ArtfSimd simd;
    bool antialias;
    bool hdr;
    bool compressTextures;
    const char* pipelineCache;
} ArtfRendererOptions;

/* Fields set by name on this stay right when the struct grows. */
ArtfRendererOptions artfRendererDefaults(int width, int height);

typedef struct ArtfRenderer ArtfRenderer;

/* NULL only when a backend was asked for by name and cannot run here (and, in this library, for AUTO
 * with no HIP device: see above). */
ArtfRenderer* artfRendererNew(const ArtfRendererOptions* options);
void artfRendererFree(ArtfRenderer* r);
ArtfBackend artfRendererBackend(const ArtfRenderer* r);
/* What draws: "HIP, AMD Instinct MI355X, 256 compute units". Borrowed. */
const char* artfRendererDescription(const ArtfRenderer* r);
void artfRendererResize(ArtfRenderer* r, int width, int height);

typedef struct ArtfRect { int x, y, w, h; } ArtfRect;
typedef struct ArtfOverlay {
    const uint32_t* pixels;
    int width, height;
    int stride; /* in pixels; 0: width */
    bool partial;
    const ArtfRect* damage;
    size_t damageCount;
} ArtfOverlay;

/* Premultiplied BGRA over the frame, from its top left. NULL: none. */
void artfRendererSetOverlay(ArtfRenderer* r, const ArtfOverlay* overlay);

bool artfRendererDraw(ArtfRenderer* r, ArtfScene* scene);
bool artfRendererReadPixels(ArtfRenderer* r, uint32_t* dst);
uint32_t* artfRendererPixels(ArtfRenderer* r);

typedef struct ArtfFrameStats {
    double cpuMs;
    double cullMs;
    double geometryMs;
    double rasterMs;          /* HIP: the tone map, on the GPU's clock */
    double lensMs;            /* HIP: the trace and the bloom, on the GPU's clock */
    uint32_t instancesDrawn;
    uint32_t trianglesBinned;
    uint32_t tilesSkipped;
    uint32_t instancesHidden;
    uint32_t traceSamples;
    double gpuMs;             /* HIP: trace + bloom + tone map */
    uint32_t shadowMapsDrawn;
} ArtfFrameStats;

void artfRendererStats(const ArtfRenderer* r, ArtfFrameStats* out);

#ifdef __cplusplus
}
#endif

#include "artful/hip_renderer.h"

#endif
