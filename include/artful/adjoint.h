/* artful-hip — differentiable rendering of the black hole with the analytic disc (beyond the
 * reference, which has no gradients): image-loss gradients with respect to the spin, the eye's
 * position, the disc's radii, its emissivity index q and its gain, by a hand-written adjoint of the
 * tracer with checkpoints every 32 steps (artful_amd/csrc/hip/adjoint.h).
 *
 * The disc is the analytic one (ARTF_DISC_ANALYTIC: emissivity (inner/r)^q with soft edges a mass
 * wide, opacity min(thickness, 1)); the description must have stars 0, velocity 0 and an explicit
 * discInner (0 would tie the inner edge to the spin through the last stable orbit). */
#ifndef ARTFUL_ADJOINT_H
#define ARTFUL_ADJOINT_H

#include "artful/renderer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ArtfHoleGradient {
    double spin;
    double position[3];  /* the camera's */
    double discInner, discOuter, emissivity, gain;
    /* dL/d the frame's fields the rays read: forward[3], right[3], up[3], metric[4], place[4], er[3],
     * etheta[3], ephi[3]; and the spin where the rays read it directly */
    double frame[26];
    double spinDirect;
} ArtfHoleGradient;

typedef struct ArtfHoleAdjoint ArtfHoleAdjoint;

/* NULL when no HIP device can run it (logged). samples: rays a pixel on a side, as the description's. */
ArtfHoleAdjoint* artfHoleAdjointNew(int width, int height, int device);
void artfHoleAdjointFree(ArtfHoleAdjoint* a);
/* The analytic disc's HDR for this description and camera; hdr: NULL or 3 host planes (R, G, B) of
 * width*height floats. tracing: its precision, steps, escape radius and q (disc is taken as analytic). */
bool artfHoleAdjointRender(ArtfHoleAdjoint* a, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, const ArtfHoleTracing* tracing, float* hdr);
/* dL/d parameters, given seed = dL/d(HDR), 3 host planes; hdr (NULL or 3 planes) gets the image the
 * adjoint's own forward pass saw. FP64 tracing gives the FP64 adjoint; anything else the FP32 one. */
bool artfHoleAdjointGradient(ArtfHoleAdjoint* a, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, const ArtfHoleTracing* tracing,
                             const float* seed, float* hdr, ArtfHoleGradient* out);
/* GPU milliseconds of the last render's trace and of the last gradient's adjoint kernel. */
void artfHoleAdjointTimes(const ArtfHoleAdjoint* a, double* forwardMs, double* adjointMs);

#ifdef __cplusplus
}
#endif

#endif
