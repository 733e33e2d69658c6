/* A driver of this project's over the reference library (Artful 0.4.0), built only into
 * oracle/_ref/libartful_ref.so by oracle/ref/Makefile. It exposes, as plain C calls for ctypes, the
 * pieces of the reference's black-hole path that golden vectors are taken from:
 *   - scene side: artfSceneSetBlackHole (src/scene/hole.c:242-271), its map and fluid;
 *   - per frame: artfHoleFrame (src/scene/hole.c:273-371);
 *   - the tracer: the `lens` slot of a kernel table (src/software/softwareInternal.h:165-182), called
 *     per 64x64 tile as drawHole does (src/software/software.c:410-418);
 *   - a whole software frame: artfRendererDraw + artfRendererReadPixels (src/render/renderer.c:68-73).
 * Test infrastructure only. */
#include "artful/artful.h"
#include "scene/sceneInternal.h"
#include "software/softwareInternal.h"
#include "render/rendererInternal.h"

#include <stdlib.h>
#include <string.h>

void* refSceneNew(const ArtfBlackHoleDesc* desc, const ArtfCamera* camera) {
    ArtfScene* s = artfSceneNew();
    artfSceneSetBlackHole(s, desc);
    artfSceneSetCamera(s, camera);
    return s;
}

void refSceneSet(void* s, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera) {
    artfSceneSetBlackHole((ArtfScene*)s, desc);
    if (camera) artfSceneSetCamera((ArtfScene*)s, camera);
}

void refSceneFree(void* s) { artfSceneFree((ArtfScene*)s); }

/* The desc as the scene keeps it, after its clamps. */
void refSceneDesc(const void* s, ArtfBlackHoleDesc* out) { *out = ((const ArtfScene*)s)->hole.desc; }

const float* refMap(const void* s) { return ((const ArtfScene*)s)->hole.map; }
const float* refFluid(const void* s) { return artfFluidField(((const ArtfScene*)s)->hole.fluid); }
double refFluidTime(const void* s) { return artfFluidTime(((const ArtfScene*)s)->hole.fluid); }
unsigned long long refMapVersion(const void* s) { return ((const ArtfScene*)s)->hole.mapVersion; }
unsigned long long refFluidVersion(const void* s) { return ((const ArtfScene*)s)->hole.fluidVersion; }
int refMapFloats(void) { return ARTF_HOLE_MAP_FLOATS; }
int refHoleFrameBytes(void) { return (int)sizeof(ArtfHoleFrame); }

uint32_t refFrame(const void* s, int width, int height, void* out) {
    uint32_t bloomPixels = 0;
    artfHoleFrame((const ArtfScene*)s, width, height, (ArtfHoleFrame*)out, &bloomPixels);
    return bloomPixels;
}

float refIsco(float spin) { return artfHoleIsco(spin); }
float refOrbitSpeed(float spin, float r) { return artfBlackHoleOrbitSpeed(spin, r); }
void refDefaults(ArtfBlackHoleDesc* out) { *out = artfBlackHoleDefaults(); }

void* refFluidNew(float inner, float outer, float spin, double time) { return artfFluidNew(inner, outer, spin, time); }
int refFluidAdvance(void* f, double time) { return artfFluidAdvance((ArtfFluid*)f, time); }
const float* refFluidField(const void* f) { return artfFluidField((const ArtfFluid*)f); }
double refFluidFieldTime(const void* f) { return artfFluidTime((const ArtfFluid*)f); }
void refFluidFree(void* f) { artfFluidFree((ArtfFluid*)f); }

/* The tracer alone, every tile, one thread: HDR into three dense planes of width * height floats.
 * isa: 0 portable, 1 SSE4.1 (2 x 128), 2 AVX2. frame: an ArtfHoleFrame as given (NULL: the scene's own). */
void refLens(const void* scene, int width, int height, int isa, const void* frame, float* outR, float* outG, float* outB) {
    const ArtfScene* s = scene;
    ArtfSoftFrame f;
    memset(&f, 0, sizeof f);
    f.width = width;
    f.height = height;
    f.tilesX = (width + ARTF_TILE - 1) / ARTF_TILE;
    f.tilesY = (height + ARTF_TILE - 1) / ARTF_TILE;
    uint32_t bloomPixels = 0;
    if (frame) memcpy(&f.holeData, frame, sizeof f.holeData);
    else artfHoleFrame(s, width, height, &f.holeData, &bloomPixels);
    f.holeMap = s->hole.map;
    f.holeFluid = artfFluidField(s->hole.fluid);
    f.hdrStride = (width + 7) & ~7;
    for (int c = 0; c < 3; ++c) f.hdr[c] = artfAlignedMalloc((size_t)f.hdrStride * height * 4 + 64, 64);
    const ArtfKernels* k = isa == 2 ? &artfKernelsAvx2 : isa == 1 ? &artfKernelsSse41 : &artfKernelsPortable;
    for (int t = 0; t < f.tilesX * f.tilesY; ++t) k->lens(&f, (uint32_t)t);
    float* outs[3] = {outR, outG, outB};
    for (int c = 0; c < 3; ++c) {
        for (int y = 0; y < height; ++y) memcpy(outs[c] + (size_t)y * width, f.hdr[c] + (size_t)y * f.hdrStride, (size_t)width * 4);
        artfAlignedFree(f.hdr[c]);
    }
}

/* A whole software frame, BGRA8. simd: an ArtfSimd value. Returns lensMs of the draw. */
double refDraw(void* scene, int width, int height, int simd, int threads, uint32_t* out) {
    ArtfRendererOptions o = artfRendererDefaults(width, height);
    o.backend = ARTF_BACKEND_SOFTWARE;
    o.threads = threads;
    o.simd = (ArtfSimd)simd;
    ArtfRenderer* r = artfRendererNew(&o);
    artfRendererDraw(r, (ArtfScene*)scene);
    artfRendererReadPixels(r, out);
    ArtfFrameStats st;
    artfRendererStats(r, &st);
    artfRendererFree(r);
    return st.lensMs;
}

/* A renderer kept across frames, for timing: new, draw n frames, stats, free. */
void* refRendererNew(int width, int height, int simd, int threads) {
    ArtfRendererOptions o = artfRendererDefaults(width, height);
    o.backend = ARTF_BACKEND_SOFTWARE;
    o.threads = threads;
    o.simd = (ArtfSimd)simd;
    return artfRendererNew(&o);
}
int refRendererDraw(void* r, void* scene) { return artfRendererDraw((ArtfRenderer*)r, (ArtfScene*)scene); }
void refRendererRead(void* r, uint32_t* out) { artfRendererReadPixels((ArtfRenderer*)r, out); }
void refRendererStats(const void* r, double* lensMs, double* rasterMs, double* cpuMs) {
    ArtfFrameStats st;
    artfRendererStats((const ArtfRenderer*)r, &st);
    *lensMs = st.lensMs;
    *rasterMs = st.rasterMs;
    *cpuMs = st.cpuMs;
}
void refRendererFree(void* r) { artfRendererFree((ArtfRenderer*)r); }
const char* refRendererDescription(const void* r) { return artfRendererDescription((const ArtfRenderer*)r); }

/* The layouts a drop-in must match: sizeof of each public struct (field "") and the offset of each of
 * its fields, as the reference's headers lay them out (include/artful/renderer.h, scene.h, core.h;
 * src/scene/sceneInternal.h for the frame's uniform block). */
#include <stddef.h>
#define SIZE(T) {#T, "", sizeof(T)}
#define AT(T, f) {#T, #f, offsetof(T, f)}
static const struct { const char* type; const char* field; size_t at; } layouts[] = {
    SIZE(ArtfRendererOptions), AT(ArtfRendererOptions, backend), AT(ArtfRendererOptions, surface), AT(ArtfRendererOptions, width),
    AT(ArtfRendererOptions, height), AT(ArtfRendererOptions, noVsync), AT(ArtfRendererOptions, threads), AT(ArtfRendererOptions, simd),
    AT(ArtfRendererOptions, antialias), AT(ArtfRendererOptions, hdr), AT(ArtfRendererOptions, compressTextures),
    AT(ArtfRendererOptions, pipelineCache),
    SIZE(ArtfSurface), AT(ArtfSurface, kind), AT(ArtfSurface, a), AT(ArtfSurface, b),
    SIZE(ArtfFrameStats), AT(ArtfFrameStats, cpuMs), AT(ArtfFrameStats, cullMs), AT(ArtfFrameStats, geometryMs), AT(ArtfFrameStats, rasterMs),
    AT(ArtfFrameStats, lensMs), AT(ArtfFrameStats, instancesDrawn), AT(ArtfFrameStats, trianglesBinned), AT(ArtfFrameStats, tilesSkipped),
    AT(ArtfFrameStats, instancesHidden), AT(ArtfFrameStats, traceSamples), AT(ArtfFrameStats, gpuMs), AT(ArtfFrameStats, shadowMapsDrawn),
    SIZE(ArtfOverlay), AT(ArtfOverlay, pixels), AT(ArtfOverlay, width), AT(ArtfOverlay, height), AT(ArtfOverlay, stride),
    AT(ArtfOverlay, partial), AT(ArtfOverlay, damage), AT(ArtfOverlay, damageCount),
    SIZE(ArtfRect), AT(ArtfRect, x), AT(ArtfRect, y), AT(ArtfRect, w), AT(ArtfRect, h),
    SIZE(ArtfVec3), SIZE(ArtfCamera), AT(ArtfCamera, position), AT(ArtfCamera, target), AT(ArtfCamera, up), AT(ArtfCamera, fovY),
    AT(ArtfCamera, near_),
    SIZE(ArtfBlackHoleDesc), AT(ArtfBlackHoleDesc, position), AT(ArtfBlackHoleDesc, axis), AT(ArtfBlackHoleDesc, mass),
    AT(ArtfBlackHoleDesc, spin), AT(ArtfBlackHoleDesc, discInner), AT(ArtfBlackHoleDesc, discOuter), AT(ArtfBlackHoleDesc, discHeight),
    AT(ArtfBlackHoleDesc, discThickness), AT(ArtfBlackHoleDesc, discGain), AT(ArtfBlackHoleDesc, doppler), AT(ArtfBlackHoleDesc, stars),
    AT(ArtfBlackHoleDesc, bloom), AT(ArtfBlackHoleDesc, exposure), AT(ArtfBlackHoleDesc, time), AT(ArtfBlackHoleDesc, shutter),
    AT(ArtfBlackHoleDesc, velocity), AT(ArtfBlackHoleDesc, samples),
    SIZE(ArtfHoleFrame), AT(ArtfHoleFrame, er), AT(ArtfHoleFrame, etheta), AT(ArtfHoleFrame, ephi), AT(ArtfHoleFrame, right),
    AT(ArtfHoleFrame, up), AT(ArtfHoleFrame, forward), AT(ArtfHoleFrame, metric), AT(ArtfHoleFrame, place), AT(ArtfHoleFrame, spin),
    AT(ArtfHoleFrame, flow), AT(ArtfHoleFrame, boost), AT(ArtfHoleFrame, disc), AT(ArtfHoleFrame, look), AT(ArtfHoleFrame, image),
    AT(ArtfHoleFrame, size), AT(ArtfHoleFrame, bloom),
};
int refLayoutCount(void) { return (int)(sizeof layouts / sizeof layouts[0]); }
const char* refLayoutType(int i) { return layouts[i].type; }
const char* refLayoutField(int i) { return layouts[i].field; }
long refLayoutAt(int i) { return (long)layouts[i].at; }
/* enum values a drop-in must keep */
int refEnumValue(int which) {
    switch (which) {
    case 0: return ARTF_BACKEND_METAL;
    case 1: return ARTF_SURFACE_METAL;
    case 2: return ARTF_SIMD_256;
    default: return -1;
    }
}
