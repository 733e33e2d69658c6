/* artful-hip — read access to a scene's black-hole data, for tests, tools and other backends (the
 * reference's backends read these through src/scene/sceneInternal.h:411-420, 616-628; this library
 * exposes them as plain calls). All returned pointers are borrowed from the scene. */
#ifndef ARTFUL_INSPECT_H
#define ARTFUL_INSPECT_H

#include "artful/hole.h"
#include "artful/scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The hole's description after artfSceneSetBlackHole's clamps; false when the scene has no hole. */
bool artfSceneHoleDesc(const ArtfScene* s, ArtfBlackHoleDesc* out);
/* ARTF_HOLE_MAP_FLOATS floats, and the version that changes when they do; NULL before any hole, and
 * when the map is in device memory (artfSceneSetDiscOnDevice). */
const float* artfSceneHoleMap(const ArtfScene* s, uint64_t* version);
/* ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT floats, their version, and the moment they are of; NULL (version and
 * moment still given) when the fluid is stepped on a device (artfSceneSetDiscOnDevice). */
const float* artfSceneHoleFluid(const ArtfScene* s, uint64_t* version, double* time);
/* The frame's uniforms at this size (levels: 0 for the reference's bloom rule, else at most that many);
 * returns the floats each bloom plane needs. */
uint32_t artfSceneHoleFrame(const ArtfScene* s, int width, int height, int levels, ArtfHoleFrame* out);
/* The same geometry in doubles. */
void artfSceneHoleFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out);
/* The frame's uniforms, FP32 and FP64, straight from a description (clamped as a scene would) and a
 * camera, without a scene (no map or fluid is built; the fluid's moment is taken as 0). */
void artfHoleFrameFor(const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, int width, int height, ArtfHoleFrame* out, ArtfHoleFrame64* out64);
/* Artful's last stable orbit (Bardeen-Press-Teukolsky). */
float artfHoleIsco(float spin);

/* The disc's fluid on its own (src/scene/fluid.c's calls): new at a moment, advanced in fixed steps
 * (true when its field changed), its field and moment. The caller owns what New returns. */
typedef struct ArtfFluid ArtfFluid;
ArtfFluid* artfFluidNew(float inner, float outer, float spin, double time);
void artfFluidFree(ArtfFluid* f);
bool artfFluidAdvance(ArtfFluid* f, double time);
const float* artfFluidField(const ArtfFluid* f);
double artfFluidTime(const ArtfFluid* f);

#ifdef __cplusplus
}
#endif

#endif
