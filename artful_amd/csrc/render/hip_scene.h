/* artful-hip — what the HIP renderer reads of a scene, and the calls through which it reads it.
 *
 * The renderer (hip_renderer.c) includes no scene internals. Each build that links it supplies these
 * few calls for its own scene code: this library's in hip_scene.c, over its scene; a build of the
 * reference with this backend added (its src/hip, beside src/vulkan) in route_b/hip_scene_reference.c,
 * over the reference's src/scene/sceneInternal.h (ArtfHole, sceneInternal.h:411-418, and artfHoleFrame,
 * :627). So the same hip_renderer.c compiles against either tree.
 *
 * Frames are passed as the struct tags both trees give them: struct ArtfHoleFrame is the 368-byte
 * uniform block of the reference's sceneInternal.h:302-319 (this library's include/artful/hole.h has
 * the same fields in the same order), struct ArtfHoleFrame64 the same geometry in doubles. */
#ifndef ARTFUL_HIP_SCENE_H
#define ARTFUL_HIP_SCENE_H

#include <stdbool.h>
#include <stdint.h>

typedef struct ArtfScene ArtfScene;
struct ArtfHoleFrame;
struct ArtfHoleFrame64;
struct ArtfRendererOptions;

/* The value ARTF_BACKEND_HIP has: the one after the reference's AUTO, VULKAN, SOFTWARE and METAL. */
#define ARTF_HIP_BACKEND_VALUE 4

/* A scene as the renderer sees it for one draw. */
typedef struct ArtfHipSceneView {
    uint64_t id;          /* no two scenes of a process share one */
    bool hole;            /* a black hole is set */
    uint32_t background;  /* BGRA of a frame with no hole: the scene's background, premultiplied by its alpha */
    const float* map;     /* host, ARTF_HOLE_MAP_FLOATS; changes only with mapVersion */
    uint64_t mapVersion;
    const float* fluid;   /* host, ARTF_FLUID_WIDTH x ARTF_FLUID_HEIGHT; changes only with fluidVersion */
    uint64_t fluidVersion;
    /* The disc's data kept on a HIP device by the scene (this library's artfSceneSetDiscOnDevice), or
     * NULL / -1: the map there, and the fluid with the calls that read its field in order. */
    const float* deviceMap;
    void* deviceFluid;
    int discDevice;
    const float* (*deviceFluidField)(void* fluid);
    bool (*deviceFluidRead)(void* fluid, float* field);     /* the current field to the host, once written */
    bool (*deviceFluidReadOn)(void* fluid, void* stream);   /* the stream's next work after the field's write */
    bool (*deviceFluidReadDone)(void* fluid, void* stream); /* the next write of the field after the stream's work */
} ArtfHipSceneView;

void artfHipSceneView(const ArtfScene* s, ArtfHipSceneView* out);
/* The frame's uniforms, as the build's artfHoleFrame fills them (the reference's bloom-level rule);
 * bloomPixels: what the levels take. */
void artfHipSceneFrame(const ArtfScene* s, int width, int height, struct ArtfHoleFrame* out, uint32_t* bloomPixels);
/* The same geometry worked out in doubles from the scene's (clamped) description and camera. */
void artfHipSceneFrame64(const ArtfScene* s, int width, int height, struct ArtfHoleFrame64* out);

/* The renderer: NULL when no HIP device can run here (logged). artfHipNew takes device 0; it is what
 * the backend dispatch (renderer.c's artfRendererNew) calls for ARTF_BACKEND_HIP and AUTO. */
struct ArtfRenderer* artfHipNew(const struct ArtfRendererOptions* options);
struct ArtfRenderer* artfHipRendererNew(const struct ArtfRendererOptions* options, int device);

/* For tests: the number of device calls that succeed before the device is reported lost, when not
 * below 0 (as the reference's artfVulkanSubmitsLeft, src/render/rendererInternal.h:27-30). */
extern int artfHipCallsLeft;

#endif
