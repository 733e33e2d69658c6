/* artful-hip — drawing a scene's black hole on an MI355X (gfx950).
 *
 * The renderer API of Artful 0.4.0 (include/artful/renderer.h:24-133) with one more backend,
 * ARTF_BACKEND_HIP, appended after the reference's values so they keep their numbers. The calls
 * keep their contracts: artfRendererNew returns NULL only when a backend asked for by name cannot
 * run; artfRendererDraw returns false when the device is lost (free the renderer, keep the scene);
 * artfRendererReadPixels writes width*height BGRA8, alpha 255; artfRendererPixels is NULL for a GPU.
 *
 * Beyond the reference: artfRendererReadHdr (the traced light before bloom, which the reference keeps
 * private) and artfRendererSetHoleTracing (precision, disc model and step law). */
#ifndef ARTFUL_RENDERER_H
#define ARTFUL_RENDERER_H

#include "artful/hole.h"
#include "artful/scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ArtfBackend {
    ARTF_BACKEND_AUTO,
    ARTF_BACKEND_VULKAN,   /* not in this library */
    ARTF_BACKEND_SOFTWARE, /* not in this library */
    ARTF_BACKEND_METAL,    /* not in this library */
    ARTF_BACKEND_HIP       /* hand-written HIP kernels on an AMD Instinct GPU */
} ArtfBackend;

const char* artfBackendName(ArtfBackend b);

typedef enum ArtfSurfaceKind { ARTF_SURFACE_NONE } ArtfSurfaceKind;
typedef struct ArtfSurface { ArtfSurfaceKind kind; void* a; void* b; } ArtfSurface;
typedef enum ArtfSimd { ARTF_SIMD_AUTO, ARTF_SIMD_PORTABLE, ARTF_SIMD_128, ARTF_SIMD_256 } ArtfSimd;

typedef struct ArtfRendererOptions {
    ArtfBackend backend; /* AUTO: HIP */
    ArtfSurface surface; /* frames stay in memory */
    int width, height;
    bool noVsync;
    int threads;         /* host threads for the disc's tables and fluid: 0 as many as allowed, at most 16 */
    ArtfSimd simd;
    bool antialias;
    bool hdr;
    bool compressTextures;
    const char* pipelineCache;
    int device;          /* HIP: which GPU (beyond the reference) */
} ArtfRendererOptions;

ArtfRendererOptions artfRendererDefaults(int width, int height);

typedef struct ArtfRenderer ArtfRenderer;

ArtfRenderer* artfRendererNew(const ArtfRendererOptions* options);
void artfRendererFree(ArtfRenderer* r);
ArtfBackend artfRendererBackend(const ArtfRenderer* r);
const char* artfRendererDescription(const ArtfRenderer* r);
void artfRendererResize(ArtfRenderer* r, int width, int height);

typedef struct ArtfRect { int x, y, w, h; } ArtfRect;
typedef struct ArtfOverlay {
    const uint32_t* pixels;
    int width, height;
    int stride;
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

/* ---- beyond the reference ---- */

/* The last frame's light before bloom: three planes (R, G, B) of width*height floats. */
bool artfRendererReadHdr(ArtfRenderer* r, float* red, float* green, float* blue);

typedef struct ArtfHoleTracing {
    int precision;      /* ArtfPrecision */
    int disc;           /* ArtfDiscModel */
    float emissivity;   /* q of the analytic disc */
    int maxSteps;       /* 0: 192, the reference's */
    float escapeRadius; /* 0: a ray is away where 1/r crosses 0, the reference's rule */
    uint32_t jitter;    /* 0: rays on the reference's stratified grid; else a seed for jittered rays */
    int bloomLevels;    /* 0: the reference's rule (while both sides exceed 3, up to 8) */
    int reserved;
} ArtfHoleTracing;

ArtfHoleTracing artfHoleTracingDefaults(void);
void artfRendererSetHoleTracing(ArtfRenderer* r, const ArtfHoleTracing* tracing);

/* The last frame's hole, stage by stage on the GPU's clock, and the tracer's work. */
typedef struct ArtfHoleStats {
    double traceMs, resolveMs, bloomMs, toneMs, uploadMs;
    uint64_t rays;      /* width * height * samples^2 */
    uint64_t steps;     /* Nystrom steps, summed over rays */
    uint64_t laneSlots; /* lanes x steps the waves spent stepping: steps / laneSlots is the active-lane fraction */
    int bloomLevels;
} ArtfHoleStats;

bool artfRendererHoleStats(const ArtfRenderer* r, ArtfHoleStats* out);

#ifdef __cplusplus
}
#endif

#endif
