/* artful-hip — the thin C ABI over the HIP kernels (gfx950). Plain C: POD arguments, device pointers
 * as plain pointers, streams as void*, an int status (0, else a HIP error number) and nothing thrown.
 * The C host backend (artful_amd/csrc/render/hip_renderer.c) calls these; so may any FFI (ctypes,
 * cgo, JNI) that wants the kernels without the renderer.
 *
 * Each entry replaces a piece of Artful 0.4.0's black-hole path, named beside it. */
#ifndef ARTFUL_HIP_H
#define ARTFUL_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "artful/core.h"
#include "artful/hole.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device, memory, streams, events (what src/vulkan/vulkan.c does through Vulkan) ---- */

int artfHipDeviceCount(void);
int artfHipSetDevice(int device);
/* The device's name and compute units into name (size bytes); 0 on success. */
int artfHipDeviceInfo(int device, char* name, size_t size, int* computeUnits);
void* artfHipMalloc(size_t bytes); /* NULL when the device has no room */
void artfHipFree(void* p);
int artfHipUpload(void* dst, const void* src, size_t bytes, void* stream);
int artfHipDownload(void* dst, const void* src, size_t bytes, void* stream);
/* Device to host without waiting: dst must be pinned (artfHipHostMalloc) and is ready once the stream
 * has been synchronised. */
int artfHipDownloadAsync(void* dst, const void* src, size_t bytes, void* stream);
/* Pinned host memory, coherent with the device (NULL if none is left), and its release: what a copy or a
 * kernel's system-scope store puts there is seen by the host once the stream has been synchronised. */
void* artfHipHostMalloc(size_t bytes);
void artfHipHostFree(void* p);
int artfHipMemset(void* dst, int value, size_t bytes, void* stream);
/* Device to device. */
int artfHipCopy(void* dst, const void* src, size_t bytes, void* stream);
void* artfHipStreamCreate(void);
void artfHipStreamDestroy(void* stream);
int artfHipStreamSync(void* stream);
/* Every stream of the current device done. */
int artfHipDeviceSync(void);
void* artfHipEventCreate(void);
void artfHipEventDestroy(void* event);
int artfHipEventRecord(void* event, void* stream);
/* An event that orders one stream after another and nothing more: no timing, a device-scope release. */
void* artfHipEventCreateOrder(void);
/* The stream's later work waits for the event's last recording. */
int artfHipStreamWait(void* stream, void* event);
/* 1 when the event's last recording has completed (or it was never recorded), 0 while it has not, -1 on
 * an error. */
int artfHipEventDone(void* event);
/* Milliseconds between two recorded events (both must have completed). */
float artfHipEventMs(void* start, void* stop);
/* The last error the runtime saw, as text. */
const char* artfHipErrorText(int status);

/* ---- the tracer: replaces lensTile + lensTrace (src/software/lens.inc:99-383), which the
 * reference calls per 64x64 tile from drawHole (src/software/software.c:410-418), and the Vulkan
 * shader artfShaderHoleTrace (src/vulkan/shaders.c:1957-2026) dispatched by traceHole
 * (src/vulkan/vulkan.c:1758-1783). ---- */

#define ARTF_HIP_COUNTER_SLOTS 1024

typedef struct ArtfHipTraceArgs {
    /* from artfHoleFrame: FP32 uniforms. artfHipTrace, artfHipResolve, artfHipBloom, artfHipToneMap,
     * artfHipDrawFrame and artfHipAdjoint take frames of at most 65535 rows and fewer than 2^31 rays
     * (side^2 * width * height), and refuse others with hipErrorInvalidValue, launching nothing. */
    ArtfHoleFrame frame;
    ArtfHoleFrame64 frame64; /* from artfHoleFrame64: read by the MIXED and FP64 tracers */
    const float* map;        /* device, ARTF_HOLE_MAP_FLOATS */
    const float* fluid;      /* device, ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT */
    /* device planes R, G, B, width * height each: the HDR image, each pixel the mean of its side^2 rays,
     * added in the order lensTile adds them and scaled by 1/side^2 (lens.inc:364-380). (artfHipAdjoint
     * writes each ray's light instead: see ArtfHipAdjointArgs.) */
    float* out[3];
    /* device, 2 * ARTF_HIP_COUNTER_SLOTS words, added to: the rays' steps in the first half, and the
     * lane-slots their waves spent stepping in the second (steps / slots is the stepping loop's
     * active-lane fraction); the totals are the sums of each half. Spread over slots so that waves do
     * not queue on one address. NULL: not counted. */
    unsigned long long* counters;
    int precision;      /* ArtfPrecision (artful/renderer.h) */
    int maxSteps;       /* 0: 192 */
    float escapeRadius; /* 0: where 1/r crosses 0 */
    int disc;           /* ArtfDiscModel */
    float emissivity;   /* the analytic disc's q */
    uint32_t jitter;    /* 0: the stratified grid; else a seed */
    int rowBegin, rowEnd; /* rows [rowBegin, rowEnd) of the image; rowEnd 0: to the last row */
    /* NULL, or a word (device or artfHipHostMalloc memory) the launch's first workgroup sets to the GPU's
     * clock (artfHipWallClockHz) as it starts */
    unsigned long long* startStamp;
    /* NULL, or a second word set with startStamp, at the same instant (a caller that queues several
     * frames gives it to the first, to time the run of them) */
    unsigned long long* spanStamp;
} ArtfHipTraceArgs;

int artfHipTrace(const ArtfHipTraceArgs* args, void* stream);

/* The mean of each pixel's side^2 rays, summed in their order and scaled by 1/side^2, as lensTile
 * does (lens.inc:364-380): for planes of each ray's light (ray k of pixel p at k * width * height + p),
 * as a batch's views and the adjoint write them; hdr: width * height each. */
int artfHipResolve(const ArtfHoleFrame* frame, float* const rays[3], float* const hdr[3], void* stream);

/* The pixels of one wave's tile at side^2 rays a pixel, the pixel's rays in adjacent lanes: 8x8 pixels at
 * one ray, 4x4 at 2x2 rays, 7x1 at 3x3, 2x2 at 4x4, 2x1 at 5x5, one pixel at 6x6 to 8x8 (with 3, 5, 6 or
 * 7 rays a side some of the 64 lanes trace nothing). The frame's tracer and the batches take the same. */
#define ARTF_HIP_TILE_W(side) ((side) == 1 ? 8 : (side) == 2 ? 4 : (side) == 3 ? 7 : (side) == 4 || (side) == 5 ? 2 : 1)
#define ARTF_HIP_TILE_H(side) ((side) == 1 ? 8 : (side) == 2 ? 4 : (side) == 4 ? 2 : 1)

/* ---- batches of views (not in the reference, which draws one view a frame): many views of any
 * size, spin, eye and disc, traced by one launch: a workgroup of one wave for each tile, or persistent
 * waves that take tiles of rays from a queue and give idle lanes new rays of the view they are on (a
 * wave holds one view's rays at a time). Each view's rays are those artfHipTrace would trace for its
 * frame, bit for bit in FP32. ---- */

typedef struct ArtfHipView {
    ArtfHoleFrame frame;     /* artfHoleFrame's FP32 uniforms for this view */
    ArtfHoleFrame64 frame64; /* read by the MIXED and FP64 tracers */
    float* hdr[3];           /* device planes of width * height: each pixel's mean, as artfHipTrace's out[] */
    /* the queued kernel's output: hdr at one ray a pixel; else device planes of each ray's light (ray k of
     * pixel p at k * width * height + p) for artfHipResolve, which it cannot sum in a wave whose lanes it
     * refills; NULL where no queued launch will run */
    float* rays[3];
    int32_t maxSteps;        /* at least 1 */
    float escapeRadius;      /* 0: where 1/r crosses 0 */
    int32_t disc;            /* ArtfDiscModel */
    float emissivity;        /* the analytic disc's q */
    uint32_t jitter;         /* 0: the stratified grid; else a seed */
    /* tiles of ARTF_HIP_TILE_W(side) x ARTF_HIP_TILE_H(side) pixels across and down the view, rounded up */
    uint32_t tilesX, tilesY;
    uint32_t rayCount;       /* width * height * side^2 */
} ArtfHipView;

/* A view's ring tables in ArtfHipBatchArgs.rings: ring speed and heat (ARTF_HOLE_MAP_HEIGHT x 2, as
 * the map's part at ARTF_HOLE_ROWS_AT), then glow per ring (ARTF_HOLE_MAP_HEIGHT, as at
 * ARTF_HOLE_GLOW_AT). */
#define ARTF_HIP_RING_FLOATS (ARTF_HOLE_MAP_HEIGHT * 3)
#define ARTF_HIP_STAMP_SLOTS 16
/* ArtfHipBatchArgs.stamps: the launch's start at word 0, then a row of slots for each k of
 * ARTF_HIP_STAMP_SLOTS, the views side by side in it, the rows 512 bytes apart or more: view v's slots
 * lie in 16 different cache lines, so the atomics of its waves ending together do not all queue on one */
#define ARTF_HIP_STAMP_ROW(views) (((uint32_t)(views) + 63u) & ~63u)
#define ARTF_HIP_STAMP_AT(views, view, k) (1u + (uint32_t)(k) * ARTF_HIP_STAMP_ROW(views) + (uint32_t)(view))
#define ARTF_HIP_STAMP_WORDS(views) (1u + ARTF_HIP_STAMP_SLOTS * ARTF_HIP_STAMP_ROW(views))
#define ARTF_HIP_QUEUES 8
#define ARTF_HIP_QUEUE_STRIDE 32

/* Whether a frame's spin lets the FP32 steps take the drag's quotient -a u (2 - xi a u) / max(1 - 2u + a^2 u^2,
 * 1e-3) by the IEEE division's own steps without its scaling and fix-up (the same quotient, two instructions
 * fewer): a spin of 0 (no drag) or of at least 2^-50 in size. artfHipTrace and the frames choose by it
 * themselves; a batch says so in ArtfHipBatchArgs.dragInRange. */
ARTF_INLINE int artfHipDragInRange(float spin) { return spin == 0.f || spin >= 0x1p-50f || spin <= -0x1p-50f; }

/* schedule.workgroups for a launch of a workgroup of one wave for each tile, with no queue */
#define ARTF_HIP_WORKGROUP_A_TILE (-1)

/* How the persistent launch spends the device. */
typedef struct ArtfHipSchedule {
    int32_t workgroups;    /* of 256 lanes (four independent waves), 1..65536; or ARTF_HIP_WORKGROUP_A_TILE */
    int32_t chunkTiles;    /* 8x8 tiles of rays a wave takes from the queue at once, 1..64 */
    int32_t refill;        /* a wave gives new rays to its idle lanes once at least this many of its 64 are idle, 1..64 */
    int32_t stepsPerCheck; /* steps between two looks at the idle lanes, 1..16 */
} ArtfHipSchedule;

typedef struct ArtfHipBatchArgs {
    const ArtfHipView* views;  /* device, viewCount */
    const uint32_t* firstTile; /* device, viewCount + 1: view v's tiles are [firstTile[v], firstTile[v + 1]), each
                                * view's rows of tiles top to bottom; ray k of a tile is lane k of its wave */
    int32_t viewCount;
    /* the tiles this launch traces, [tileBegin, tileEnd) of [0, firstTile[viewCount]): a huge batch is
     * traced by several launches of a slice each (watchdog-sized), the queues cleared between them */
    uint32_t tileBegin, tileEnd;
    int32_t precision;         /* ArtfPrecision, the same for every view */
    const float* map;          /* device, ARTF_HOLE_MAP_FLOATS: its gas, colours, sky and dust are shared */
    const float* rings;        /* device, viewCount * ARTF_HIP_RING_FLOATS: view v's at v * ARTF_HIP_RING_FLOATS */
    /* device, viewCount * ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT: view v's fluid field at v times that; read
     * for painted views only (NULL if there are none) */
    const float* fluids;
    /* device workspace, cleared by artfHipBatchReset (or, between the slices of one batch,
     * artfHipBatchResetQueues) before each launch: ARTF_HIP_QUEUES counters, each ARTF_HIP_QUEUE_STRIDE
     * words from the last. Queue q hands out the q-th eighth of the launch's tiles to the workgroups of
     * its XCD (blockIdx % ARTF_HIP_QUEUES, the order workgroups are dealt to XCDs), which then take from
     * the other queues until all are empty. */
    uint32_t* queue;
    /* device, 2 * ARTF_HIP_COUNTER_SLOTS, as artfHipTrace's counters (lane-slots: every lane of every
     * stepping wave, idle ones included); NULL: not counted */
    unsigned long long* counters;
    /* device, ARTF_HIP_STAMP_WORDS(viewCount) words, or NULL: the launch's start (stamps[0]), then per
     * view v the latest moment a wave finished rays of it, in slots ARTF_HIP_STAMP_AT(viewCount, v, k)
     * (the view's last ray ended at the largest of its slots), in the device's wall clock
     * (artfHipWallClockHz). Spread over slots, as the counters are. */
    unsigned long long* stamps;
    ArtfHipSchedule schedule;
    /* 1 only if every view's frame.spin[0] passes artfHipDragInRange: the FP32 steps then take the drag's
     * quotient without the division's scaling (the same bits, fewer instructions); 0: by IEEE division,
     * right for any spin */
    int32_t dragInRange;
} ArtfHipBatchArgs;

/* Clears the queue, the counters and the stamps, and stamps the start, with a kernel (no
 * memset node, so that the pair can be captured in a graph and replayed). */
int artfHipBatchReset(const ArtfHipBatchArgs* args, void* stream);
/* Clears the queues alone, between two launches of one batch (the counters and stamps run on). */
int artfHipBatchResetQueues(const ArtfHipBatchArgs* args, void* stream);
int artfHipTraceViews(const ArtfHipBatchArgs* args, void* stream);
/* Workgroups of the persistent kernel for this precision one compute unit holds at once. */
int artfHipViewsOccupancy(int precision);
/* The frequency of the clock the stamps count, in Hz (0 if unknown). */
double artfHipWallClockHz(int device);

/* ---- graphs: what a stream does between Begin and End, replayed as one launch ---- */
int artfHipGraphBegin(void* stream);
/* *graph: an executable graph, to launch with artfHipGraphLaunch and free with artfHipGraphFree. */
int artfHipGraphEnd(void* stream, void** graph);
int artfHipGraphLaunch(void* graph, void* stream);
void artfHipGraphFree(void* graph);

/* ---- the disc map on the GPU: replaces buildDisc (src/scene/hole.c:136-178), bit for bit ---- */

/* Paints the map's gas (normalised), the disc's dust and the sky's dust into a device map of
 * ARTF_HOLE_MAP_FLOATS; the colour table and the per-ring rows and glow are the host's to fill
 * (artfDiscColors / artfDiscRingTables in the scene code: a few thousand floats). */
int artfHipDiscPaint(float* map, void* stream);

/* ---- the disc fluid on the GPU: replaces the host passes of artfFluidNew / artfFluidAdvance
 * (src/scene/fluid.c:96-241, 255-336), bit for bit. The host keeps the random stream and the ring
 * constants (artfFluidRingsOf, artfFluidDrawEddy in the scene code) and hands over eddies. ---- */

typedef struct ArtfHipFluid {
    /* device, ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT floats each, ring by ring: the state ... */
    float* density;
    float* radial;
    float* around;
    float* scratch[3]; /* ... a step's second buffers (density, radial, around) ... */
    float* field;      /* ... and what shows, for the tracer */
    const ArtfFluidRings* rings; /* device */
} ArtfHipFluid;

/* The starting clumps (fluid.c:286-305); velocities zero. */
int artfHipFluidSeed(const ArtfHipFluid* fluid, void* stream);
/* Eddies (device, count of them) stirred in, in their order. */
int artfHipFluidStir(const ArtfHipFluid* fluid, const ArtfFluidEddy* eddies, int count, void* stream);
/* One step (fluid.c:231-241): its ARTF_FLUID_STEP_EDDIES eddies (device), then the velocity, density,
 * shift and balance passes. */
int artfHipFluidStep(const ArtfHipFluid* fluid, const ArtfFluidEddy* eddies, void* stream);
/* field from the state (fluid.c:164-176). */
int artfHipFluidField(const ArtfHipFluid* fluid, void* stream);

/* ---- the adjoint (not in the reference): image-loss gradients of the analytic-disc tracer ---- */

typedef struct ArtfHipAdjointArgs {
    /* the ray law and buffers, as for artfHipTrace: disc must be ARTF_DISC_ANALYTIC, precision FP32 or
     * FP64, the frame's stars 0 and its eye at rest (boost w 1), maxSteps at most 2048. out[] receives
     * each ray's light: side^2 planes of width * height (ray k of pixel p at k * width * height + p; one
     * ray a pixel, the image); artfHipResolve averages them. */
    ArtfHipTraceArgs trace;
    const float* seed[3]; /* device: dL/d(HDR) of each pixel, R G B planes of width*height */
    /* device: 32 doubles per workgroup (artfHipAdjointBlocks of them), each workgroup's sum over its
     * rays of dL/d: forward[3], right[3], up[3], metric[4], place[4], er[3], etheta[3], ephi[3] (the
     * frame's fields), the spin where the rays read it directly, disc inner, outer, q, gain. */
    double* partials;
} ArtfHipAdjointArgs;

int artfHipAdjointBlocks(const ArtfHoleFrame* frame);
int artfHipAdjoint(const ArtfHipAdjointArgs* args, void* stream);

/* ---- the post passes: replace bloomDownRow/bloomUpRow (src/software/software.c:320-386, 419-426)
 * and lensFill's tone map (src/software/lens.inc:387-440), and the Vulkan shaders artfShaderBloomDown,
 * artfShaderBloomUp and artfShaderHoleBackdrop (src/vulkan/shaders.c:2058-2142, 2207-2257). ---- */

/* The bloom pyramid: frame->size[3] levels, laid out per frame->bloom[k] in three planes of
 * frame-dependent length (artfHoleFrame's bloomPixels each). */
int artfHipBloom(const ArtfHoleFrame* frame, float* const hdr[3], float* const bloom[3], void* stream);
/* HDR + level 0 of the bloom to BGRA8 (alpha 255), width * height words. */
int artfHipToneMap(const ArtfHoleFrame* frame, float* const hdr[3], float* const bloom[3], uint32_t* bgra, void* stream);

/* One frame of the hole, as the reference's drawHole and lensFill make it (src/software/software.c:409-427,
 * lens.inc:355-440): artfHipTrace (each pixel's rays summed in its lanes), then artfHipBloom and
 * artfHipToneMap, with the same bytes out, in fewer launches and with nothing else on the stream: the
 * pyramid's small levels are one workgroup's, bloom level 0's kernel sums the trace's counters into the
 * report and clears them (no fill, no copy), and each stage's kernel stamps the GPU's clock into the
 * report (no events between kernels). */
#define ARTF_HIP_REPORT_WORDS 7
typedef struct ArtfHipFrameArgs {
    /* out[]: the HDR planes, hdr[] below; whole frames (rowBegin 0, rowEnd 0 or the height); startStamp:
     * set by artfHipDrawFrame (spanStamp is kept as given). counters: zero before the first frame (each
     * frame leaves them zero), or NULL: not counted, the report's steps and lane-slots then 0. */
    ArtfHipTraceArgs trace;
    float* hdr[3];   /* device, width * height each */
    float* bloom[3]; /* device, the pyramid as artfHipBloom lays it out */
    uint32_t* bgra;  /* device, width * height words */
    /* NULL, or host memory from artfHipHostMalloc, ARTF_HIP_REPORT_WORDS words, there once the streams are
     * synchronised: [0] the frame's steps and [1] lane-slots (the sums of the counters' halves); the GPU's
     * clock when [2] the trace started, [3] bloom level 0 started, [4] the rest of the bloom started (with
     * no more levels: the tone map), [5] the tone map started,
     * [6] the tone map's last workgroup ended. Frames queued one after another share it: it holds the
     * last one's. */
    unsigned long long* report;
    /* NULL: the whole frame on the call's stream. Else the trace alone goes on the call's stream, and the
     * bloom and the tone map on this second one, after `handoff` (artfHipEventCreateOrder) is recorded
     * past the trace; `done` (the same kind, or NULL) is recorded after them. The call's stream is then
     * free for the next frame's trace while they run: the caller keeps later frames off what they still
     * read and write (the counters, the HDR and the pyramid: a second set for the next frame, its stream
     * waiting for `done` before a set is used again), and waits for both streams before reading the
     * frame back. */
    void* postStream;
    void* handoff;
    void* done;
} ArtfHipFrameArgs;

/* Frames of more than ARTF_HIP_RAYS_APART rays are traced in bands of rows of at least ARTF_HIP_RAYS_A_BAND
 * rays, a launch each, one after another on the call's stream (the reference's Vulkan backend splits a frame
 * of more than 64,000,000 rays into bands too, src/vulkan/vulkan.c:1729-1755); the same image. */
#define ARTF_HIP_RAYS_APART 64000000ull
#define ARTF_HIP_RAYS_A_BAND 64000000ull

int artfHipDrawFrame(const ArtfHipFrameArgs* args, void* stream);

/* artfHipDrawFrame for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc); it has no counterpart in the reference. The same
 * frame, bytes and report, with the trace's step counters (args->trace.counters is ignored and may be
 * NULL) in a static device array of 2 * ARTF_HIP_COUNTER_SLOTS words (16 KiB), one per GPU, held until
 * the process ends, whatever the frame's size; frames are limited as artfHipDrawFrame's are (at most
 * 65535 rows, fewer than 2^31 rays), others refused with hipErrorInvalidValue before any launch. It
 * calls no allocator. The whole frame runs on `stream` (postStream must be NULL: refused otherwise).
 * Every call on one GPU shares that GPU's array, zeroed by a kernel at each call's start: calls that
 * overlap on one GPU (other streams, other threads) race on it; calls in order on one stream do not. */
int artfHipDrawFrameStatic(const ArtfHipFrameArgs* args, void* stream);

#ifdef __cplusplus
}
#endif

#endif
