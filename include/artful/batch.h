/* artful-hip — many views of black holes traced in one launch (beyond the reference, which draws one
 * view a frame through artfRendererDraw, renderer.h). Each view has its own description (spin, disc,
 * samples, ...), camera and size; all share a precision and ray law. The views' rays go through one
 * launch, replayed from a HIP graph (the workspace reset, the trace, each view's resolve): by default a
 * workgroup of one wave for each 8x8 tile of rays, or a persistent, work-queued kernel; artfHoleBatchTune
 * picks the schedule by timing candidates on the batch itself.
 *
 * Each view's HDR equals what the HIP renderer draws for a scene with the same description and
 * camera at the same time, bit for bit in FP32. A view's disc tables (rings, and the fluid for the
 * painted disc) are its own, built when the batch is made; the painted gas, colours and sky are
 * built once for all. */
#ifndef ARTFUL_BATCH_H
#define ARTFUL_BATCH_H

#include "artful/renderer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ArtfHoleView {
    ArtfBlackHoleDesc hole; /* as for artfSceneSetBlackHole (clamped the same way) */
    ArtfCamera camera;
    int32_t width, height;
} ArtfHoleView;

/* workgroups for no persistent launch: a workgroup of one wave for each 8x8 tile of rays, handed out by
 * the hardware as the frame's tracer is (chunkTiles, refill and stepsPerCheck then do not matter) */
#define ARTF_HOLE_WORKGROUP_A_TILE (-1)

/* How the launch spends the device (the first four fields as artful_hip.h's ArtfHipSchedule). */
typedef struct ArtfHoleSchedule {
    /* 0 (the default) or below: ARTF_HOLE_WORKGROUP_A_TILE; 1..65536: that many persistent workgroups of
     * 256 lanes take tiles from queues (artfHipViewsOccupancy of them fit a compute unit at once) */
    int32_t workgroups;
    int32_t chunkTiles;    /* 8x8 tiles a wave takes from the queue at once */
    int32_t refill;        /* idle lanes of 64 that make a wave take new rays */
    int32_t stepsPerCheck; /* steps between two looks at the idle lanes */
    /* 0: all tiles in one launch; else launches of at most this many 8x8 tiles each, one after another
     * in the graph (a huge batch in watchdog-sized slices) */
    int32_t launchTiles;
} ArtfHoleSchedule;

typedef struct ArtfHoleBatchStats {
    double batchMs;     /* the last draw: the graph's launch to its end, on the GPU's clock */
    double traceMs;     /* its trace alone: the start stamp to the last view's last ray */
    uint64_t rays;      /* all views' width * height * samples^2 */
    uint64_t steps;     /* Nystrom steps, summed over rays */
    uint64_t laneSlots; /* lanes x steps the waves spent stepping, idle lanes included */
    int32_t views;
    ArtfHoleSchedule schedule; /* the one in use, zeros filled in */
} ArtfHoleBatchStats;

typedef struct ArtfHoleBatch ArtfHoleBatch;

/* NULL when no HIP device can run it, a view is unusable or memory runs out (logged). tracing: NULL
 * for artfHoleTracingDefaults (its bloomLevels is not used: a batch draws HDR). */
ArtfHoleBatch* artfHoleBatchNew(const ArtfHoleView* views, int count, const ArtfHoleTracing* tracing, int device);
void artfHoleBatchFree(ArtfHoleBatch* b);
/* Traces every view once; false if the device failed. */
bool artfHoleBatchDraw(ArtfHoleBatch* b);
/* View v's light: three planes (R, G, B) of its width*height floats, into hdr. */
bool artfHoleBatchReadHdr(ArtfHoleBatch* b, int view, float* hdr);
bool artfHoleBatchStats(const ArtfHoleBatch* b, ArtfHoleBatchStats* out);
/* Per view, the milliseconds from the last draw's start to the end of that view's last ray. */
bool artfHoleBatchLatency(const ArtfHoleBatch* b, double* ms);
/* NULL or zeros: the defaults. The next draw uses it (its graph is captured again). */
bool artfHoleBatchSetSchedule(ArtfHoleBatch* b, const ArtfHoleSchedule* schedule);
/* Times candidate schedules on this batch (a few draws each, coordinate by coordinate), keeps the
 * fastest and writes it to best (may be NULL). rounds: passes over the coordinates, 0 for 2. */
bool artfHoleBatchTune(ArtfHoleBatch* b, int rounds, ArtfHoleSchedule* best);

#ifdef __cplusplus
}
#endif

#endif
