/* artful-hip — a renderer: an object with a table of functions, as Artful 0.4.0's
 * src/render/rendererInternal.h:7-24 (ArtfRendererVtbl, struct ArtfRenderer), slot for slot and field
 * for field, under the same file name: the HIP backend (hip_renderer.c) includes it as
 * "render/rendererInternal.h", and so compiles against this library's copy or the reference's own.
 * This library has one backend; a build of the reference's tree with this backend added declares
 * artfHipNew beside artfVulkanNew (its rendererInternal.h:32-33). */
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
} ArtfRendererVtbl;

struct ArtfRenderer {
    const ArtfRendererVtbl* vt;
    ArtfBackend backend;
    /* the reference's occlusion trial (renderer.c:43-66); triangle scenes are not drawn here */
    int trialFrame;
    double trialMs[2];
};

/* NULL when no HIP device can run here (logged). */
ArtfRenderer* artfHipNew(const ArtfRendererOptions* options);

#endif
