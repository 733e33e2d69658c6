/* artful-hip — the scene's insides, as the HIP backend reads them. */
#ifndef ARTFUL_SCENE_INTERNAL_H
#define ARTFUL_SCENE_INTERNAL_H

#include "artful/hole.h"
#include "artful/scene.h"
#include "frame64.h"

typedef struct ArtfFluid ArtfFluid;

/* The hole of a scene: its clamped description, its tables and its fluid. map and fluid change
 * only with their versions, which is what a backend compares to know when to upload them again. */
typedef struct ArtfHole {
    bool on;
    ArtfBlackHoleDesc desc;
    float* map;            /* ARTF_HOLE_MAP_FLOATS on the host, or */
    float* deviceMap;      /* the same map in the memory of HIP device discDevice (>= 0) */
    uint64_t mapVersion;
    ArtfFluid* fluid;      /* the host fluid, or */
    struct ArtfDeviceFluid* deviceFluid; /* the same fluid stepped on device discDevice */
    int discDevice;
    uint64_t fluidVersion;
} ArtfHole;

/* The time of the scene's fluid, wherever it is (0 without one). */
double artfHoleFluidTime(const ArtfHole* h);

struct ArtfScene {
    uint64_t id;
    ArtfCamera camera;
    ArtfHole hole;
    ArtfVec3 background;   /* what shows without a hole (scene.c:39-40 defaults) */
    float backgroundAlpha;
};

/* ---- disc.c: the painted gas and the per-ring tables (Artful src/scene/hole.c:10-197) ---- */
void artfDiscPaint(float* map);  /* gas, dust, colours, sky: once per map */
void artfDiscRings(float* map, float inner, float outer, float spin);
/* The same tables apart from a map: rows ARTF_HOLE_MAP_HEIGHT x 2 (speed, heat), glow ARTF_HOLE_MAP_HEIGHT. */
void artfDiscRingTables(float* rows, float* glow, float inner, float outer, float spin);

/* ---- fluid.c (Artful src/scene/fluid.c) ---- */
#define ARTF_FLUID_CATCH_UP 48 /* most steps one advance takes */
#define ARTF_FLUID_WARM 60     /* steps before the first field */
#define ARTF_FLUID_SEED 0x9E3779B9u /* the random stream's first word */
ArtfFluid* artfFluidNew(float inner, float outer, float spin, double time);
/* The ring constants of a fluid on this disc; and the next eddy of a fluid's random stream (rng). */
void artfFluidRingsOf(float inner, float outer, float spin, ArtfFluidRings* out);
void artfFluidDrawEddy(uint32_t* rng, const ArtfFluidRings* k, float inner, float outer, bool small, ArtfFluidEddy* e);
void artfFluidFree(ArtfFluid* f);
bool artfFluidAdvance(ArtfFluid* f, double time);
const float* artfFluidField(const ArtfFluid* f);
double artfFluidTime(const ArtfFluid* f);

/* ---- hole.c ---- */
float artfHoleIsco(float spin);
/* The frame's uniforms (Artful src/scene/hole.c:273-371); bloomPixels: what the bloom's levels take.
 * levelsCap: 0 for the reference's rule, else at most that many bloom levels. */
void artfHoleFrame(const ArtfScene* s, int width, int height, ArtfHoleFrame* out, uint32_t* bloomPixels);
void artfHoleFrameLevels(const ArtfScene* s, int width, int height, int levelsCap, ArtfHoleFrame* out, uint32_t* bloomPixels);
/* The same geometry from the same description, worked out in doubles. */
void artfHoleFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out);
/* The same, from a description (already clamped: artfHoleClamped) and a camera rather than a scene. */
ArtfBlackHoleDesc artfHoleClamped(const ArtfBlackHoleDesc* desc);
void artfHoleFrameOf(const ArtfBlackHoleDesc* d, const ArtfCamera* c, double fluidTime, int width, int height, int levelsCap, ArtfHoleFrame* out,
                     uint32_t* bloomPixels);
/* The black-body colour table alone (ARTF_HOLE_COLORS x RGB), as the map holds it. */
void artfDiscColors(float* colors);

#endif
