/* artful-hip — a renderer: an object with a table of functions, as Artful 0.4.0's
 * src/render/rendererInternal.h:7-24 (ArtfRendererVtbl, struct ArtfRenderer). This library has one
 * backend, HIP; a build that also links the reference's software or Vulkan backend adds their tables. */
#ifndef ARTFUL_RENDERER_INTERNAL_H
#define ARTFUL_RENDERER_INTERNAL_H

#include "artful/renderer.h"

typedef struct ArtfRendererVtbl {
    const char* (*description)(const ArtfRenderer* r);
    void (*resize)(ArtfRenderer* r, int width, int height);
    void (*setOverlay)(ArtfRenderer* r, const ArtfOverlay* overlay);
    bool (*draw)(ArtfRenderer* r, ArtfScene* scene);
    bool (*readPixels)(ArtfRenderer* r, uint32_t* dst);
    uint32_t* (*pixels)(ArtfRenderer* r);
    void (*stats)(const ArtfRenderer* r, ArtfFrameStats* out);
    void (*destroy)(ArtfRenderer* r);
    /* beyond the reference */
    bool (*readHdr)(ArtfRenderer* r, float* red, float* green, float* blue);
    void (*setHoleTracing)(ArtfRenderer* r, const ArtfHoleTracing* tracing);
    bool (*holeStats)(const ArtfRenderer* r, ArtfHoleStats* out);
} ArtfRendererVtbl;

struct ArtfRenderer {
    const ArtfRendererVtbl* vt;
    ArtfBackend backend;
};

/* NULL when no HIP device can run here (logged). */
ArtfRenderer* artfHipNew(const ArtfRendererOptions* options);

/* For tests: the number of device calls that succeed before the device is reported lost, when not
 * below 0 (as the reference's artfVulkanSubmitsLeft, rendererInternal.h:27-30). */
extern int artfHipCallsLeft;

#endif
