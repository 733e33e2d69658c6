/* artful-hip — a batch of black-hole views traced by one launch (include/artful/batch.h).
 *
 * Made once, on the device (include/artful/device_disc.h): the painted map's shared tables (gas,
 * colours, sky, dust) for all views; per view its clamped description, FP32 and FP64 frames, ring
 * tables and (painted disc) fluid
 * with the view table and the views' tile offsets. Each draw replays one HIP graph: the workspace
 * reset kernel, the trace over all views (a workgroup a tile, which writes each pixel's mean, or
 * persistent and queued, which writes each ray's light), and after the queued kernel the resolve of each
 * view with several rays a pixel. The tuner times candidate schedules on the batch itself and keeps the
 * fastest. */
#include <stdio.h>
#include <string.h>

#include "../scene/scene_internal.h"
#include "artful/batch.h"
#include "artful/device_disc.h"
#include "artful_hip.h"

#define RING_FLOATS ARTF_HIP_RING_FLOATS
#define FLUID_FLOATS ((size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT)

typedef struct BatchView {
    ArtfHoleFrame frame;
    int width, height, side;
    float* hdr[3];  /* device, width * height */
    float* rays[3]; /* device, width * height * side^2 when side > 1, once a queued schedule is captured */
} BatchView;

struct ArtfHoleBatch {
    int count, device, computeUnits;
    ArtfHoleTracing tracing;
    BatchView* views;
    void* stream;
    void* events[2];
    void* side;               /* the second stream of a draw in slices */
    void* order[2];           /* fork and join of the two streams */
    float* map;               /* device, ARTF_HOLE_MAP_FLOATS */
    float* rings;             /* device, count * RING_FLOATS */
    float* fluids;            /* device, count * FLUID_FLOATS, painted disc only */
    ArtfHipView* deviceViews; /* device, count */
    ArtfHipView* hostViews;   /* the same, as uploaded */
    bool rayPlanes;           /* the views with several rays a pixel have their planes of rays */
    uint32_t* firstTile;      /* device, count + 1 */
    void* workspace;          /* device: two sets of queues, counters, stamps */
    uint32_t* queues[2];      /* in the workspace: a set of queues for each stream */
    ArtfHipBatchArgs args;
    unsigned long long* countersHost;
    unsigned long long* stampsHost;
    uint64_t rays;
    uint32_t tiles;
    ArtfHoleSchedule schedule;
    void* graph;
    double batchMs, clockHz;
    bool drawn;
};

static void freePlanes(float** p) {
    for (int c = 0; c < 3; ++c) artfHipFree(p[c]), p[c] = NULL;
}

static bool allocPlanes(float** p, size_t n) {
    for (int c = 0; c < 3; ++c)
        if (!(p[c] = artfHipMalloc(n * sizeof(float)))) return false;
    return true;
}

void artfHoleBatchFree(ArtfHoleBatch* b) {
    if (!b) return;
    if (b->stream) artfHipStreamSync(b->stream);
    artfHipGraphFree(b->graph);
    if (b->views)
        for (int v = 0; v < b->count; ++v) freePlanes(b->views[v].hdr), freePlanes(b->views[v].rays);
    artfFree(b->views);
    artfDeviceMapFree(b->map), artfHipFree(b->rings), artfHipFree(b->fluids);
    artfHipFree(b->deviceViews), artfHipFree(b->firstTile), artfHipFree(b->workspace);
    artfFree(b->hostViews);
    artfFree(b->countersHost), artfFree(b->stampsHost);
    if (b->side) artfHipStreamSync(b->side);
    for (int i = 0; i < 2; ++i) artfHipEventDestroy(b->events[i]), artfHipEventDestroy(b->order[i]);
    artfHipStreamDestroy(b->stream);
    artfHipStreamDestroy(b->side);
    artfFree(b);
}

/* The schedule with its zeros filled in and its fields in range. */
static ArtfHoleSchedule resolved(const ArtfHoleSchedule* s) {
    ArtfHoleSchedule r = s ? *s : (ArtfHoleSchedule){0};
    /* the default: a workgroup a tile, which outran every queued schedule on the batches measured */
    if (r.workgroups <= 0) r.workgroups = ARTF_HOLE_WORKGROUP_A_TILE;
    else r.workgroups = artfClampi(r.workgroups, 1, 65536);
    r.chunkTiles = r.chunkTiles <= 0 ? 2 : artfClampi(r.chunkTiles, 1, 64);
    r.refill = r.refill <= 0 ? 64 : artfClampi(r.refill, 1, 64);
    r.stepsPerCheck = r.stepsPerCheck <= 0 ? 8 : artfClampi(r.stepsPerCheck, 1, 16);
    r.launchTiles = r.launchTiles < 0 ? 0 : r.launchTiles;
    return r;
}

static ArtfHoleBatch* fail(ArtfHoleBatch* b, const char* what) {
    artfLog("artful: the batch could not %s", what);
    artfHoleBatchFree(b);
    return NULL;
}

static double largest(double a, double b) { return a > b ? a : b; }

ArtfHoleBatch* artfHoleBatchNew(const ArtfHoleView* views, int count, const ArtfHoleTracing* tracing, int device) {
    if (!views || count < 1 || count > 65536) {
        artfLog("artful: a batch needs 1 to 65536 views");
        return NULL;
    }
    ArtfHoleTracing t = tracing ? *tracing : artfHoleTracingDefaults();
    if (t.precision < ARTF_PRECISION_FP32 || t.precision > ARTF_PRECISION_FP64 || (t.disc != ARTF_DISC_PAINTED && t.disc != ARTF_DISC_ANALYTIC) ||
        t.maxSteps < 0 || t.maxSteps > 1000000 || !(t.escapeRadius >= 0.f)) {
        artfLog("artful: the batch's tracing is out of range");
        return NULL;
    }
    if (t.maxSteps == 0) t.maxSteps = 192;
    uint64_t tiles = 0;
    for (int v = 0; v < count; ++v) {
        const ArtfBlackHoleDesc d = artfHoleClamped(&views[v].hole);
        const int w = views[v].width, h = views[v].height;
        if (w < 1 || h < 1 || w > 32768 || h > 32768 || (uint64_t)w * h * d.samples * d.samples >= (1ull << 31)) {
            artfLog("artful: view %d's size %dx%d is out of range", v, w, h);
            return NULL;
        }
        tiles += (uint64_t)((w + ARTF_HIP_TILE_W(d.samples) - 1) / ARTF_HIP_TILE_W(d.samples)) * ((h + ARTF_HIP_TILE_H(d.samples) - 1) / ARTF_HIP_TILE_H(d.samples));
    }
    if (tiles > (1u << 25)) {
        artfLog("artful: a batch holds at most 2^25 tiles of 64 rays");
        return NULL;
    }
    if (device < 0 || device >= artfHipDeviceCount() || artfHipSetDevice(device)) {
        artfLog("artful: no HIP device %d for the batch", device);
        return NULL;
    }

    ArtfHoleBatch* b = ARTF_NEW(ArtfHoleBatch);
    b->count = count, b->device = device, b->tracing = t;
    b->views = artfCalloc((size_t)count, sizeof(BatchView));
    char name[256];
    if (artfHipDeviceInfo(device, name, sizeof name, &b->computeUnits) || b->computeUnits < 1) return fail(b, "read the device");
    b->clockHz = artfHipWallClockHz(device);
    b->stream = artfHipStreamCreate();
    b->side = artfHipStreamCreate();
    bool ready = b->stream != NULL && b->side != NULL;
    for (int i = 0; i < 2; ++i)
        ready = ready && (b->events[i] = artfHipEventCreate()) != NULL && (b->order[i] = artfHipEventCreateOrder()) != NULL;
    if (!ready) return fail(b, "make its stream");

    /* the shared tables, painted once on the device */
    const bool painted = t.disc == ARTF_DISC_PAINTED;
    float* rings = artfMalloc((size_t)count * RING_FLOATS * sizeof(float));
    ArtfHipView* hv = artfCalloc((size_t)count, sizeof(ArtfHipView));
    uint32_t* firstTile = artfMalloc((size_t)(count + 1) * sizeof(uint32_t));
    b->map = artfDeviceMapNew(device);
    b->rings = artfHipMalloc((size_t)count * RING_FLOATS * sizeof(float));
    b->fluids = painted ? artfHipMalloc((size_t)count * FLUID_FLOATS * sizeof(float)) : NULL;
    ready = b->map && b->rings && (b->fluids || !painted);

    uint32_t at = 0;
    for (int v = 0; ready && v < count; ++v) {
        const ArtfHoleView* given = &views[v];
        const ArtfBlackHoleDesc d = artfHoleClamped(&given->hole);
        BatchView* bv = &b->views[v];
        bv->width = given->width, bv->height = given->height, bv->side = d.samples;
        float* rows = rings + (size_t)v * RING_FLOATS;
        artfDiscRingTables(rows, rows + ARTF_HOLE_MAP_HEIGHT * 2, d.discInner, d.discOuter, d.spin);
        double fluidTime = d.time;
        if (painted) {
            /* the fluid a scene with this description would hold, made at its time as artfSceneSetBlackHole
             * does, but stepped on the device (bit for bit the host's) */
            ArtfDeviceFluid* fluid = artfDeviceFluidNew(d.discInner, d.discOuter, d.spin, given->hole.time, device);
            ready = fluid && !artfHipCopy(b->fluids + (size_t)v * FLUID_FLOATS, artfDeviceFluidField(fluid), FLUID_FLOATS * sizeof(float), b->stream) &&
                    !artfHipStreamSync(b->stream);
            if (fluid) fluidTime = artfDeviceFluidTime(fluid);
            artfDeviceFluidFree(fluid);
        }
        artfHoleFrameOf(&d, &given->camera, fluidTime, bv->width, bv->height, 1, &bv->frame, NULL);
        const size_t pixels = (size_t)bv->width * bv->height;
        ready = ready && allocPlanes(bv->hdr, pixels);
        if (!ready) break;
        ArtfHipView* h = &hv[v];
        h->frame = bv->frame;
        artfHoleFrame64Of(&d, &given->camera, bv->width, bv->height, &h->frame64);
        for (int c = 0; c < 3; ++c) h->hdr[c] = bv->hdr[c], h->rays[c] = bv->side > 1 ? NULL : bv->hdr[c];
        h->maxSteps = t.maxSteps, h->escapeRadius = t.escapeRadius, h->disc = t.disc, h->emissivity = t.emissivity, h->jitter = t.jitter;
        const int tileW = ARTF_HIP_TILE_W(bv->side), tileH = ARTF_HIP_TILE_H(bv->side);
        h->tilesX = (uint32_t)((bv->width + tileW - 1) / tileW), h->tilesY = (uint32_t)((bv->height + tileH - 1) / tileH);
        h->rayCount = (uint32_t)(pixels * bv->side * bv->side);
        firstTile[v] = at;
        at += h->tilesX * h->tilesY;
        b->rays += h->rayCount;
    }
    firstTile[count] = at;

    const size_t slots = 2 * ARTF_HIP_COUNTER_SLOTS, stamps = ARTF_HIP_STAMP_WORDS(count);
    const size_t queueBytes = (size_t)ARTF_HIP_QUEUES * ARTF_HIP_QUEUE_STRIDE * 4;
    const size_t workBytes = 2 * queueBytes + slots * 8 + stamps * 8;
    if (ready) {
        b->deviceViews = artfHipMalloc((size_t)count * sizeof(ArtfHipView));
        b->firstTile = artfHipMalloc((size_t)(count + 1) * sizeof(uint32_t));
        b->workspace = artfHipMalloc(workBytes);
        ready = b->deviceViews && b->firstTile && b->workspace && !artfHipUpload(b->rings, rings, (size_t)count * RING_FLOATS * sizeof(float), b->stream) &&
                !artfHipUpload(b->deviceViews, hv, (size_t)count * sizeof(ArtfHipView), b->stream) &&
                !artfHipUpload(b->firstTile, firstTile, (size_t)(count + 1) * sizeof(uint32_t), b->stream) && !artfHipStreamSync(b->stream);
    }
    artfHipStreamSync(b->stream); /* the uploads have read the host copies */
    artfFree(rings), artfFree(firstTile);
    b->hostViews = hv;
    b->rayPlanes = true;
    for (int v = 0; v < count; ++v) b->rayPlanes = b->rayPlanes && b->views[v].side == 1;
    if (!ready) return fail(b, "set up its device memory");

    char* w = (char*)b->workspace;
    ArtfHipBatchArgs* a = &b->args;
    a->views = b->deviceViews, a->firstTile = b->firstTile, a->viewCount = count, a->tileBegin = 0, a->tileEnd = at;
    b->tiles = at;
    a->precision = t.precision, a->map = b->map, a->rings = b->rings, a->fluids = b->fluids;
    a->dragInRange = 1;
    for (int v = 0; v < count; ++v) a->dragInRange &= artfHipDragInRange(b->views[v].frame.spin[0]);
    b->queues[0] = (uint32_t*)w, b->queues[1] = (uint32_t*)(w + queueBytes);
    a->queue = b->queues[0];
    a->counters = (unsigned long long*)(w + 2 * queueBytes);
    a->stamps = (unsigned long long*)(w + 2 * queueBytes + slots * 8);
    b->countersHost = artfMalloc(slots * 8);
    b->stampsHost = artfCalloc(stamps, 8);
    b->schedule = resolved(NULL);
    return b;
}

/* The planes of rays the queued kernel writes for the views with several rays a pixel (it cannot sum a
 * pixel's rays in a wave whose lanes it refills), made the first time a queued schedule is captured, and
 * the view table uploaded again with them. */
static bool makeRayPlanes(ArtfHoleBatch* b) {
    if (b->rayPlanes) return true;
    for (int v = 0; v < b->count; ++v) {
        BatchView* bv = &b->views[v];
        if (bv->side == 1 || bv->rays[0]) continue;
        if (!allocPlanes(bv->rays, (size_t)bv->width * bv->height * bv->side * bv->side)) {
            freePlanes(bv->rays);
            artfLog("artful: the batch's planes of rays do not fit on the device");
            return false;
        }
        for (int c = 0; c < 3; ++c) b->hostViews[v].rays[c] = bv->rays[c];
    }
    if (artfHipUpload(b->deviceViews, b->hostViews, (size_t)b->count * sizeof(ArtfHipView), b->stream) || artfHipStreamSync(b->stream)) return false;
    b->rayPlanes = true;
    return true;
}

/* The graph of one draw: the reset, the trace over all views and, after the queued kernel, the resolve
 * of each view with several rays a pixel. In slices of launchTiles tiles, the slices take turns on two
 * streams, each with its own set of queues (cleared before each of its slices, for the queued kernel): a
 * slice's last waves, as its launch drains, run beside the next slice's first, instead of the device
 * idling between launches. */
static bool capture(ArtfHoleBatch* b) {
    const bool queued = b->schedule.workgroups != ARTF_HOLE_WORKGROUP_A_TILE;
    if (queued && !makeRayPlanes(b)) return false;
    ArtfHipBatchArgs* a = &b->args;
    a->schedule.workgroups = b->schedule.workgroups, a->schedule.chunkTiles = b->schedule.chunkTiles;
    a->schedule.refill = b->schedule.refill, a->schedule.stepsPerCheck = b->schedule.stepsPerCheck;
    const uint32_t slice = b->schedule.launchTiles > 0 ? (uint32_t)b->schedule.launchTiles : b->tiles;
    const bool split = b->tiles > slice;
    void* streams[2] = {b->stream, b->side};
    if (artfHipGraphBegin(b->stream)) return false;
    ArtfHipBatchArgs x = *a;
    if (b->tracing.uncounted) x.counters = NULL; /* the steps not counted: the stats' steps and lane-slots read 0 */
    x.queue = b->queues[0], x.tileBegin = 0, x.tileEnd = slice < b->tiles ? slice : b->tiles;
    int status = artfHipBatchReset(&x, b->stream); /* queues of the first set, counters, stamps */
    if (!status && split) { /* the second stream joins once the reset is done */
        status = artfHipEventRecord(b->order[0], b->stream);
        if (!status) status = artfHipStreamWait(b->side, b->order[0]);
    }
    uint32_t k = 0;
    for (uint32_t begin = 0; !status && begin < b->tiles; begin += slice, ++k) {
        x.queue = b->queues[k & 1];
        x.tileBegin = begin, x.tileEnd = b->tiles - begin > slice ? begin + slice : b->tiles;
        if (k > 0) status = artfHipBatchResetQueues(&x, streams[k & 1]);
        if (!status) status = artfHipTraceViews(&x, streams[k & 1]);
    }
    if (!status && split) { /* and rejoins before the resolves */
        status = artfHipEventRecord(b->order[1], b->side);
        if (!status) status = artfHipStreamWait(b->stream, b->order[1]);
    }
    for (int v = 0; queued && !status && v < b->count; ++v) /* the tile form wrote each pixel's mean itself */
        if (b->views[v].side > 1) status = artfHipResolve(&b->views[v].frame, b->views[v].rays, b->views[v].hdr, b->stream);
    void* graph = NULL;
    const int ended = artfHipGraphEnd(b->stream, &graph);
    if (status || ended) {
        artfLog("artful: the batch's graph could not be captured: %s", artfHipErrorText(status ? status : ended));
        artfHipGraphFree(graph);
        return false;
    }
    b->graph = graph;
    return true;
}

bool artfHoleBatchDraw(ArtfHoleBatch* b) {
    if (!b || artfHipSetDevice(b->device)) return false;
    if (!b->graph && !capture(b)) return false;
    if (artfHipEventRecord(b->events[0], b->stream) || artfHipGraphLaunch(b->graph, b->stream) || artfHipEventRecord(b->events[1], b->stream) ||
        artfHipStreamSync(b->stream)) {
        artfLog("artful: the batch's draw failed on the device");
        return false;
    }
    b->batchMs = artfHipEventMs(b->events[0], b->events[1]);
    if (artfHipDownload(b->countersHost, b->args.counters, 2 * ARTF_HIP_COUNTER_SLOTS * 8, b->stream) ||
        artfHipDownload(b->stampsHost, b->args.stamps, (size_t)ARTF_HIP_STAMP_WORDS(b->count) * 8, b->stream))
        return false;
    b->drawn = true;
    return true;
}

bool artfHoleBatchReadHdr(ArtfHoleBatch* b, int view, float* hdr) {
    if (!b || !hdr || view < 0 || view >= b->count || !b->drawn) return false;
    const BatchView* bv = &b->views[view];
    const size_t pixels = (size_t)bv->width * bv->height;
    for (int c = 0; c < 3; ++c)
        if (artfHipDownload(hdr + c * pixels, bv->hdr[c], pixels * sizeof(float), b->stream)) return false;
    return true;
}

static double stampMs(const ArtfHoleBatch* b, int view) {
    const unsigned long long start = b->stampsHost[0];
    unsigned long long end = 0;
    for (int k = 0; k < ARTF_HIP_STAMP_SLOTS; ++k) {
        const unsigned long long t = b->stampsHost[ARTF_HIP_STAMP_AT(b->count, view, k)];
        end = t > end ? t : end;
    }
    if (b->clockHz <= 0 || end < start) return -1.0;
    return (double)(end - start) * 1e3 / b->clockHz;
}

bool artfHoleBatchStats(const ArtfHoleBatch* b, ArtfHoleBatchStats* out) {
    if (!b || !out) return false;
    memset(out, 0, sizeof *out);
    out->views = b->count, out->rays = b->rays, out->schedule = b->schedule;
    if (!b->drawn) return true;
    out->batchMs = b->batchMs;
    for (int i = 0; !b->tracing.uncounted && i < ARTF_HIP_COUNTER_SLOTS; ++i)
        out->steps += b->countersHost[i], out->laneSlots += b->countersHost[ARTF_HIP_COUNTER_SLOTS + i];
    for (int v = 0; v < b->count; ++v) out->traceMs = largest(out->traceMs, stampMs(b, v));
    return true;
}

bool artfHoleBatchLatency(const ArtfHoleBatch* b, double* ms) {
    if (!b || !ms || !b->drawn) return false;
    for (int v = 0; v < b->count; ++v) ms[v] = stampMs(b, v);
    return true;
}

bool artfHoleBatchSetSchedule(ArtfHoleBatch* b, const ArtfHoleSchedule* schedule) {
    if (!b) return false;
    const ArtfHoleSchedule s = resolved(schedule);
    if (b->graph && !memcmp(&s, &b->schedule, sizeof s)) return true;
    b->schedule = s;
    artfHipGraphFree(b->graph);
    b->graph = NULL;
    return true;
}

/* The best of a few draws' times under a schedule; negative if a draw failed. */
static double timed(ArtfHoleBatch* b, const ArtfHoleSchedule* s) {
    if (!artfHoleBatchSetSchedule(b, s) || !artfHoleBatchDraw(b)) return -1.0;
    double best = 1e30;
    for (int i = 0; i < 3; ++i) {
        if (!artfHoleBatchDraw(b)) return -1.0;
        best = b->batchMs < best ? b->batchMs : best;
    }
    return best;
}

bool artfHoleBatchTune(ArtfHoleBatch* b, int rounds, ArtfHoleSchedule* best) {
    if (!b) return false;
    if (rounds <= 0) rounds = 2;
    const int occupancy = artfMaxi(artfHipViewsOccupancy(b->tracing.precision), 1), cus = b->computeUnits;
    int32_t groups[9];
    int groupCount = 0;
    for (int k = 1; k <= occupancy && groupCount < 7; ++k) groups[groupCount++] = cus * k;
    groups[groupCount++] = cus * occupancy * 2; /* more than fit: the rest wait for a slot */
    groups[groupCount++] = ARTF_HOLE_WORKGROUP_A_TILE;
    const int32_t chunks[] = {1, 2, 4, 8, 16};
    const int32_t refills[] = {4, 8, 16, 24, 32, 48, 64};
    const int32_t strides[] = {1, 2, 4, 8};
    const int32_t* choices[4] = {groups, chunks, refills, strides};
    const int counts[4] = {groupCount, 5, 7, 4};

    ArtfHoleSchedule now = b->schedule;
    double nowMs = timed(b, &now);
    if (nowMs < 0) return false;
    for (int round = 0; round < rounds; ++round) {
        for (int coord = 0; coord < 4; ++coord) {
            if (coord > 0 && now.workgroups == ARTF_HOLE_WORKGROUP_A_TILE) break; /* the queue's fields do not apply */
            for (int i = 0; i < counts[coord]; ++i) {
                ArtfHoleSchedule trial = now;
                int32_t* field = coord == 0 ? &trial.workgroups : coord == 1 ? &trial.chunkTiles : coord == 2 ? &trial.refill : &trial.stepsPerCheck;
                if (*field == choices[coord][i]) continue;
                *field = choices[coord][i];
                const double ms = timed(b, &trial);
                if (ms < 0) return false;
                if (ms < nowMs) now = trial, nowMs = ms;
            }
        }
    }
    if (!artfHoleBatchSetSchedule(b, &now)) return false;
    if (best) *best = b->schedule;
    return true;
}
