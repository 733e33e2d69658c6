/* artful-hip — the choice of a backend and the calls every backend answers (Artful 0.4.0
 * src/render/renderer.c:4-78, same contracts). */
#include "renderer_internal.h"

const char* artfBackendName(ArtfBackend b) {
    switch (b) {
    case ARTF_BACKEND_VULKAN: return "Vulkan";
    case ARTF_BACKEND_SOFTWARE: return "software";
    case ARTF_BACKEND_METAL: return "Metal";
    case ARTF_BACKEND_HIP: return "HIP";
    default: return "auto";
    }
}

/* NULL only when a backend asked for by name cannot run here; AUTO takes HIP. */
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

ArtfHoleTracing artfHoleTracingDefaults(void) {
    ArtfHoleTracing t;
    memset(&t, 0, sizeof t);
    t.precision = ARTF_PRECISION_FP32;
    t.disc = ARTF_DISC_PAINTED;
    t.emissivity = 3.f;
    return t;
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
bool artfRendererReadHdr(ArtfRenderer* r, float* red, float* green, float* blue) { return r->vt->readHdr(r, red, green, blue); }
void artfRendererSetHoleTracing(ArtfRenderer* r, const ArtfHoleTracing* tracing) { r->vt->setHoleTracing(r, tracing); }
bool artfRendererHoleStats(const ArtfRenderer* r, ArtfHoleStats* out) {
    memset(out, 0, sizeof *out);
    return r->vt->holeStats(r, out);
}
