/* artful-hip — the HIP renderer's view of this library's scenes (the calls of hip_scene.h). */
#include "hip_scene.h"

#include "../scene/scene_internal.h"
#include "artful/device_disc.h"

static const float* deviceField(void* f) { return artfDeviceFluidField((ArtfDeviceFluid*)f); }
static bool deviceRead(void* f, float* field) { return artfDeviceFluidRead((const ArtfDeviceFluid*)f, field); }
static bool deviceReadOn(void* f, void* stream) { return artfDeviceFluidReadOn((ArtfDeviceFluid*)f, stream); }
static bool deviceReadDone(void* f, void* stream) { return artfDeviceFluidReadDone((ArtfDeviceFluid*)f, stream); }

/* A byte of a colour as the reference packs its background (software.c:597-598). */
static uint32_t byteOf(float v) { return (uint32_t)(artfClampf(v, 0, 1) * 255.f + .5f); }

void artfHipSceneView(const ArtfScene* s, ArtfHipSceneView* out) {
    memset(out, 0, sizeof *out);
    const ArtfHole* h = &s->hole;
    out->id = s->id;
    out->hole = h->on;
    /* the background premultiplied by its alpha, then packed (scene.c:1322-1324, software.c:597-598) */
    const float alpha = s->backgroundAlpha;
    const ArtfVec3 bg = artfVec3Scale(s->background, alpha);
    out->background = byteOf(alpha) << 24 | byteOf(bg.x) << 16 | byteOf(bg.y) << 8 | byteOf(bg.z);
    out->map = h->map;
    out->mapVersion = h->mapVersion;
    out->fluid = h->fluid ? artfFluidField(h->fluid) : NULL;
    out->fluidVersion = h->fluidVersion;
    out->deviceMap = h->deviceMap;
    out->deviceFluid = h->deviceFluid;
    out->discDevice = h->discDevice;
    out->deviceFluidField = deviceField;
    out->deviceFluidRead = deviceRead;
    out->deviceFluidReadOn = deviceReadOn;
    out->deviceFluidReadDone = deviceReadDone;
}

void artfHipSceneFrame(const ArtfScene* s, int width, int height, ArtfHoleFrame* out, uint32_t* bloomPixels) {
    artfHoleFrame(s, width, height, out, bloomPixels);
}

void artfHipSceneFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out) { artfHoleFrame64(s, width, height, out); }
