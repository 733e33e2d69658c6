/* artful-hip — the HIP renderer's view of the reference's own scenes: route B, this backend built into
 * Artful 0.4.0's tree (as its src/hip, beside src/vulkan) with the reference's scene code left as it is.
 *
 * Compiled against the reference's headers (its include/ and src/ ahead of this repository's on the
 * include path), it reads what the Vulkan backend reads of a scene (vulkan.c:1694-1727, 2566-2584):
 * scene->id, scene->hole.on, .map, .mapVersion, .fluid (through artfFluidField), .fluidVersion
 * (src/scene/sceneInternal.h:411-418, 422-461), the background and its alpha (:454, 457), and each
 * frame's uniforms from the reference's artfHoleFrame (:627). The FP64 frame comes from this
 * repository's frame64.c, which reads only the clamped description and the camera. The reference keeps
 * its disc on the host, so no device map or device fluid is offered. Not part of this library's own
 * build (hip_scene.c is its counterpart there); oracle/ref/Makefile builds it, target routeb. */
#include "scene/sceneInternal.h"

#include "../../scene/frame64.h"
#include "../hip_scene.h"

/* A byte of a colour as the reference packs its background (software.c:597-598). */
static uint32_t byteOf(float v) { return (uint32_t)(artfClampf(v, 0, 1) * 255.f + .5f); }

void artfHipSceneView(const ArtfScene* s, ArtfHipSceneView* out) {
    memset(out, 0, sizeof *out);
    const ArtfHole* h = &s->hole;
    out->id = s->id;
    out->hole = h->on;
    const float alpha = s->backgroundAlpha;
    const ArtfVec3 bg = artfVec3Scale(s->background, alpha);
    out->background = byteOf(alpha) << 24 | byteOf(bg.x) << 16 | byteOf(bg.y) << 8 | byteOf(bg.z);
    out->map = h->map;
    out->mapVersion = h->mapVersion;
    out->fluid = h->fluid ? artfFluidField(h->fluid) : NULL;
    out->fluidVersion = h->fluidVersion;
    out->discDevice = -1;
}

void artfHipSceneFrame(const ArtfScene* s, int width, int height, struct ArtfHoleFrame* out, uint32_t* bloomPixels) {
    artfHoleFrame(s, width, height, out, bloomPixels);
}

void artfHipSceneFrame64(const ArtfScene* s, int width, int height, struct ArtfHoleFrame64* out) {
    artfHoleFrame64Of(&s->hole.desc, &s->camera, width, height, out);
}
