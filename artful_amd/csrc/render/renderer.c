/* artful-hip — the choice of a backend and the calls every backend answers (Artful 0.4.0
 * src/render/renderer.c:4-78, same contracts). This library has one backend, HIP: AUTO and HIP take
 * it, and with no HIP device both give NULL (logged); asking by name for the reference's others gives
 * NULL too. */
#include "render/rendererInternal.h"

const char* artfBackendName(ArtfBackend b) {
    switch (b) {
    case ARTF_BACKEND_VULKAN: return "Vulkan";
    case ARTF_BACKEND_SOFTWARE: return "software";
    case ARTF_BACKEND_METAL: return "Metal";
    case ARTF_BACKEND_HIP: return "HIP";
    default: return "auto";
    }
}

ArtfRenderer* artfRendererNew(const ArtfRendererOptions* o) {
    ArtfRenderer* r = NULL;
    if (o->backend == ARTF_BACKEND_AUTO || o->backend == ARTF_BACKEND_HIP) r = artfHipNew(o);
    if (!r) artfLog("artful: no %s backend here", artfBackendName(o->backend));
    return r;
}

ArtfRendererOptions artfRendererDefaults(int width, int height) {
    ArtfRendererOptions o;
    memset(&o, 0, sizeof o);
    o.width = width;
    o.height = height;
    return o;
}

void artfRendererFree(ArtfRenderer* r) {
    if (r) r->vt->destroy(r);
}
ArtfBackend artfRendererBackend(const ArtfRenderer* r) { return r->backend; }
const char* artfRendererDescription(const ArtfRenderer* r) { return r->vt->description(r); }
void artfRendererResize(ArtfRenderer* r, int width, int height) { r->vt->resize(r, width, height); }
void artfRendererSetOverlay(ArtfRenderer* r, const ArtfOverlay* overlay) { r->vt->setOverlay(r, overlay); }
bool artfRendererDraw(ArtfRenderer* r, ArtfScene* scene) { return r->vt->draw(r, scene); }
bool artfRendererReadPixels(ArtfRenderer* r, uint32_t* dst) { return r->vt->readPixels(r, dst); }
uint32_t* artfRendererPixels(ArtfRenderer* r) { return r->vt->pixels(r); }
void artfRendererStats(const ArtfRenderer* r, ArtfFrameStats* out) {
    memset(out, 0, sizeof *out);
    r->vt->stats(r, out);
}
