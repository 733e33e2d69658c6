/* artful-hip — the scene, as far as the black-hole path needs it: a camera, a background and a hole.
 *
 * Same names, layouts and meanings as Artful 0.4.0's include/artful/scene.h:312-322 (ArtfCamera,
 * artfSceneSetCamera), :343-347 (artfSceneSetBackground, artfSceneSetBackgroundAlpha) and :465-500 (ArtfBlackHoleDesc, artfBlackHoleDefaults,
 * artfBlackHoleOrbitSpeed, artfSceneSetBlackHole); artfSceneNew/artfSceneFree as in
 * src/scene/scene.c:27-52. Triangle scenes (meshes, materials, lights, glTF) are not part of this
 * library. A scene is used from one thread, and does not change while a renderer draws it. */
#ifndef ARTFUL_SCENE_H
#define ARTFUL_SCENE_H

#include "artful/core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ArtfScene ArtfScene;

/* The caller owns the result: artfSceneFree. */
ArtfScene* artfSceneNew(void);
void artfSceneFree(ArtfScene* s);

typedef struct ArtfCamera {
    ArtfVec3 position;
    ArtfVec3 target;
    ArtfVec3 up;
    float fovY;  /* radians */
    float near_; /* unused by the hole */
} ArtfCamera;

void artfSceneSetCamera(ArtfScene* s, const ArtfCamera* camera);

/* What shows where nothing is drawn, 0..1 for red, green, blue (scene.h:343-344): here, a frame with no
 * hole. (.05, .06, .08) to start with. */
void artfSceneSetBackground(ArtfScene* s, ArtfVec3 color);
/* How much of the background there is, 0..1 (1 to start with): a frame with no hole has this alpha and
 * its colour premultiplied by it; a black hole fills the frame all the same (scene.h:345-347). */
void artfSceneSetBackgroundAlpha(ArtfScene* s, float alpha);

/* A Kerr black hole with a glowing disc, lengths in masses (scene.h:465-490). 92 bytes. */
typedef struct ArtfBlackHoleDesc {
    ArtfVec3 position;
    ArtfVec3 axis;       /* what the disc turns around; zero: +y */
    float mass;          /* in the world's lengths */
    float spin;          /* -1..1 of the most a hole can have (clamped to +-.998) */
    float discInner;     /* masses; 0: the last stable orbit for this spin */
    float discOuter;
    float discHeight;    /* half depth per unit of radius */
    float discThickness; /* how much of the light behind it the disc stops */
    float discGain;      /* brightness */
    float doppler;       /* 0: no Doppler shift; 1: the full relativistic shift and beaming */
    float stars;         /* 0: none */
    float bloom;         /* the lens's veil */
    float exposure;
    float time;          /* masses; the gas flows forward only */
    float shutter;       /* exposure time across a pixel's rays: motion blur */
    ArtfVec3 velocity;   /* the eye's, as a fraction of light's */
    int samples;         /* rays a pixel on a side, 1..8 */
} ArtfBlackHoleDesc;

/* Spin .6, disc from the last stable orbit to 34, axis (0,-1,0), gain 2.1, stars .4, bloom 4.5. */
ArtfBlackHoleDesc artfBlackHoleDefaults(void);
/* A circular orbit's speed at r against one at rest in the turning space there, at most .99. */
float artfBlackHoleOrbitSpeed(float spin, float r);
/* NULL removes the hole. Otherwise: clamps, rebuilds the per-ring tables and the fluid when the
 * spin or the disc changes, else moves the fluid on to desc->time. */
void artfSceneSetBlackHole(ArtfScene* s, const ArtfBlackHoleDesc* desc);

#ifdef __cplusplus
}
#endif

#endif
