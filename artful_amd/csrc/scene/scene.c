/* artful-hip — a scene: a camera, a background and, perhaps, a black hole (Artful 0.4.0
 * src/scene/scene.c:27-78, 941-942, 1285, as far as the hole's path goes). */
#include "scene_internal.h"

#include "artful/device_disc.h"

#include <stdatomic.h>

#include "artful/inspect.h"

ArtfScene* artfSceneNew(void) {
    static atomic_ullong lastId;
    ArtfScene* s = ARTF_NEW(ArtfScene);
    s->id = atomic_fetch_add(&lastId, 1) + 1;
    s->camera.fovY = 1.f;
    s->camera.near_ = .1f;
    s->camera.up = artfVec3(0, 1, 0);
    s->camera.position = artfVec3(0, 0, 1);
    s->hole.discDevice = -1;
    s->background = artfVec3(.05f, .06f, .08f); /* the reference's scene.c:39-40 */
    s->backgroundAlpha = 1.f;
    return s;
}

void artfSceneFree(ArtfScene* s) {
    if (!s) return;
    artfFree(s->hole.map);
    artfFluidFree(s->hole.fluid);
    artfDeviceFluidFree(s->hole.deviceFluid);
    artfDeviceMapFree(s->hole.deviceMap);
    artfFree(s);
}

void artfSceneSetCamera(ArtfScene* s, const ArtfCamera* camera) { s->camera = *camera; }
void artfSceneSetBackground(ArtfScene* s, ArtfVec3 color) { s->background = color; }
void artfSceneSetBackgroundAlpha(ArtfScene* s, float alpha) { s->backgroundAlpha = artfClampf(alpha, 0, 1); }

bool artfSceneHoleDesc(const ArtfScene* s, ArtfBlackHoleDesc* out) {
    if (!s->hole.on) return false;
    *out = s->hole.desc;
    return true;
}

const float* artfSceneHoleMap(const ArtfScene* s, uint64_t* version) {
    if (version) *version = s->hole.mapVersion;
    return s->hole.map;
}

const float* artfSceneHoleFluid(const ArtfScene* s, uint64_t* version, double* time) {
    if (version) *version = s->hole.fluidVersion;
    if (time) *time = artfHoleFluidTime(&s->hole);
    return s->hole.fluid ? artfFluidField(s->hole.fluid) : NULL;
}

uint32_t artfSceneHoleFrame(const ArtfScene* s, int width, int height, int levels, ArtfHoleFrame* out) {
    uint32_t bloomPixels = 0;
    artfHoleFrameLevels(s, width, height, levels, out, &bloomPixels);
    return bloomPixels;
}

void artfSceneHoleFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out) { artfHoleFrame64(s, width, height, out); }

void artfHoleFrameFor(const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, int width, int height, ArtfHoleFrame* out, ArtfHoleFrame64* out64) {
    const ArtfBlackHoleDesc d = artfHoleClamped(desc);
    if (out) artfHoleFrameOf(&d, camera, 0.0, width, height, 0, out, NULL);
    if (out64) artfHoleFrame64Of(&d, camera, width, height, out64);
}
