/* artful-hip — what the HIP backend adds to the renderer API (none of it is in Artful 0.4.0, whose
 * API has no way to ask for these): a renderer on a chosen device, the traced light before bloom
 * (the reference keeps HDR private, src/software/softwareInternal.h:115-117), the precision and ray
 * law of the tracer, and each stage's GPU time with the tracer's step counts.
 *
 * The calls take any ArtfRenderer; for one of another backend they do nothing and say so (false, or
 * a log line). Built into the reference's tree, a program includes this header beside the
 * reference's artful/renderer.h. */
#ifndef ARTFUL_HIP_RENDERER_H
#define ARTFUL_HIP_RENDERER_H

#include "artful/hole.h"
#include "artful/renderer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A HIP renderer on HIP device `device` (artfRendererNew's HIP and AUTO take device 0). NULL, logged,
 * when that device cannot run it. */
ArtfRenderer* artfHipRendererNew(const ArtfRendererOptions* options, int device);

/* The last frame's light before bloom: three planes (R, G, B) of width*height floats. */
bool artfRendererReadHdr(ArtfRenderer* r, float* red, float* green, float* blue);

typedef struct ArtfHoleTracing {
    int precision;      /* ArtfPrecision */
    int disc;           /* ArtfDiscModel */
    float emissivity;   /* q of the analytic disc */
    int maxSteps;       /* 0: 192, the reference's */
    float escapeRadius; /* 0: a ray is away where 1/r crosses 0, the reference's rule */
    uint32_t jitter;    /* 0: rays on the reference's stratified grid; else a seed for jittered rays */
    int bloomLevels;    /* 0: the reference's rule (while both sides exceed 3, up to 8) */
    /* 0: each frame's Nystrom steps counted into ArtfHoleStats (steps, laneSlots); 1: not counted (those
     * read 0), and the trace skips the two wave reductions and two atomics a wave the counts take, 1.5% of
     * a 1080p trace. A batch (artful/batch.h) made with it counts nothing either (its stats' steps and
     * laneSlots read 0). */
    int uncounted;
} ArtfHoleTracing;

ArtfHoleTracing artfHoleTracingDefaults(void);
void artfRendererSetHoleTracing(ArtfRenderer* r, const ArtfHoleTracing* tracing);

/* The last frame's hole, stage by stage on the GPU's clock, and the tracer's work; and, for the
 * frames drawn since the stats were last read (the renderer queues frames and waits for them only
 * when something is read), the GPU-clock span from the first one's trace start to the last one's
 * tone-map end. */
typedef struct ArtfHoleStats {
    double traceMs, resolveMs, bloomMs, toneMs, uploadMs;
    uint64_t rays;      /* width * height * samples^2 */
    uint64_t steps;     /* Nystrom steps, summed over rays */
    uint64_t laneSlots; /* lanes x steps the waves spent stepping: steps / laneSlots is the active-lane fraction */
    int bloomLevels;
    uint32_t frames;    /* frames drawn since the stats were last read */
    double spanMs;      /* their GPU-clock span, first trace start to last tone-map end */
} ArtfHoleStats;

bool artfRendererHoleStats(const ArtfRenderer* r, ArtfHoleStats* out);

#ifdef __cplusplus
}
#endif

#endif
