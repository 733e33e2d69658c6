/* artful-hip — the disc's data on the GPU (include/artful/device_disc.h).
 *
 * The map: gas, dust and sky painted by hip/disc.hip; the colour table and the ring rows and glow are
 * the host's own code (artfDiscColors, artfDiscRingTables), a few thousand floats, uploaded.
 * The fluid: the host's fluid (scene/fluid.c) with its passes moved to the kernels of hip/fluid.hip.
 * The random stream, the clock and the ring constants stay here and are the host fluid's own code
 * (artfFluidRingsOf, artfFluidDrawEddy), so the eddies handed to the device are the very ones the host
 * would stir. */
#include <stdio.h>
#include <string.h>

#include "../scene/scene_internal.h"
#include "artful/device_disc.h"
#include "artful_hip.h"

#define CELLS ((size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT)
#define FIRST_STIRS 200 /* large eddies before the warm-up steps (fluid.c:306) */

float* artfDeviceMapNew(int device) {
    if (device < 0 || device >= artfHipDeviceCount() || artfHipSetDevice(device)) {
        artfLog("artful: no HIP device %d for the disc map", device);
        return NULL;
    }
    const size_t bytes = (size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float);
    float* map = artfHipMalloc(bytes);
    float colors[ARTF_HOLE_COLORS * 3];
    artfDiscColors(colors);
    if (!map || artfHipMemset(map, 0, bytes, NULL) || artfHipDiscPaint(map, NULL) ||
        artfHipUpload(map + ARTF_HOLE_COLORS_AT, colors, sizeof colors, NULL) || artfHipStreamSync(NULL)) {
        artfLog("artful: the disc map could not be painted on the device");
        artfHipFree(map);
        return NULL;
    }
    return map;
}

bool artfDeviceMapRings(float* map, float inner, float outer, float spin, int device) {
    float rows[ARTF_HOLE_MAP_HEIGHT * 2], glow[ARTF_HOLE_MAP_HEIGHT];
    artfDiscRingTables(rows, glow, inner, outer, spin);
    /* a renderer may still be tracing through the old rings: the device's queued work first */
    return !artfHipSetDevice(device) && !artfHipDeviceSync() && !artfHipUpload(map + ARTF_HOLE_ROWS_AT, rows, sizeof rows, NULL) &&
           !artfHipUpload(map + ARTF_HOLE_GLOW_AT, glow, sizeof glow, NULL) && !artfHipStreamSync(NULL);
}

void artfDeviceMapFree(float* map) {
    if (!map) return;
    artfHipDeviceSync();
    artfHipFree(map);
}

#define READERS 8 /* marked reads of a field buffer kept before the oldest is waited for at once */

struct ArtfDeviceFluid {
    int device;
    void* stream;
    float inner, outer;
    ArtfFluidRings rings;
    uint32_t rng;
    double time;
    void* block; /* device: six fields, then two field buffers, then the ring constants */
    ArtfHipFluid f;
    ArtfFluidEddy* eddies; /* device, room for eddyRoom */
    int eddyRoom;
    ArtfFluidEddy* eddiesHost;
    /* the field, in two buffers taken in turn by queued advances: the current one, the event after
     * each buffer's last write, and the marked reads of each since */
    float* fields[2];
    int current;
    void* written[2];
    void* reads[2][READERS];
    int readCount[2];
};

void artfDeviceFluidFree(ArtfDeviceFluid* f) {
    if (!f) return;
    /* a renderer may still be tracing through the field */
    if (!artfHipSetDevice(f->device)) artfHipDeviceSync();
    artfHipFree(f->block);
    artfHipFree(f->eddies);
    artfFree(f->eddiesHost);
    for (int b = 0; b < 2; ++b) {
        artfHipEventDestroy(f->written[b]);
        for (int k = 0; k < READERS; ++k) artfHipEventDestroy(f->reads[b][k]);
    }
    artfHipStreamDestroy(f->stream);
    artfFree(f);
}

/* Room for n eddies on both sides. */
static bool eddyRoom(ArtfDeviceFluid* f, int n) {
    if (n <= f->eddyRoom) return true;
    artfHipStreamSync(f->stream);
    artfHipFree(f->eddies);
    artfFree(f->eddiesHost);
    f->eddies = artfHipMalloc((size_t)n * sizeof(ArtfFluidEddy));
    f->eddiesHost = artfMalloc((size_t)n * sizeof(ArtfFluidEddy));
    f->eddyRoom = f->eddies ? n : 0;
    return f->eddies != NULL;
}

/* steps steps: their eddies drawn in order on the host, sent at once, then the passes. eddiesHost is
 * pageable memory: a copy from it has read it when hipMemcpyAsync returns (the runtime stages pageable
 * sources), so the next call may draw into it while this one's passes are still queued. */
static bool run(ArtfDeviceFluid* f, int steps) {
    const int n = steps * ARTF_FLUID_STEP_EDDIES;
    if (!eddyRoom(f, n)) return false;
    for (int s = 0; s < steps; ++s)
        for (int e = 0; e < ARTF_FLUID_STEP_EDDIES; ++e)
            artfFluidDrawEddy(&f->rng, &f->rings, f->inner, f->outer, e >= 2, &f->eddiesHost[s * ARTF_FLUID_STEP_EDDIES + e]);
    if (artfHipUpload(f->eddies, f->eddiesHost, (size_t)n * sizeof(ArtfFluidEddy), f->stream)) return false;
    for (int s = 0; s < steps; ++s) {
        if (artfHipFluidStep(&f->f, f->eddies + s * ARTF_FLUID_STEP_EDDIES, f->stream)) return false;
        f->time += ARTF_FLUID_DT;
    }
    return true;
}

/* The field of the fluid as stepped, into the other buffer once its marked reads have ended; waited
 * for here when `wait`. */
static bool finish(ArtfDeviceFluid* f, bool wait) {
    const int b = f->current ^ 1;
    for (int k = 0; k < f->readCount[b]; ++k)
        if (artfHipStreamWait(f->stream, f->reads[b][k])) return false;
    f->readCount[b] = 0;
    f->f.field = f->fields[b];
    if (artfHipFluidField(&f->f, f->stream) || artfHipEventRecord(f->written[b], f->stream)) return false;
    f->current = b;
    return !wait || !artfHipStreamSync(f->stream);
}

ArtfDeviceFluid* artfDeviceFluidNew(float inner, float outer, float spin, double time, int device) {
    if (device < 0 || device >= artfHipDeviceCount() || artfHipSetDevice(device)) {
        artfLog("artful: no HIP device %d for the fluid", device);
        return NULL;
    }
    ArtfDeviceFluid* f = ARTF_NEW(ArtfDeviceFluid);
    f->device = device, f->inner = inner, f->outer = outer, f->rng = ARTF_FLUID_SEED, f->time = time;
    artfFluidRingsOf(inner, outer, spin, &f->rings);
    f->stream = artfHipStreamCreate();
    f->block = artfHipMalloc(8 * CELLS * sizeof(float) + sizeof(ArtfFluidRings));
    bool ready = f->stream && f->block;
    for (int b = 0; b < 2; ++b) {
        ready = ready && (f->written[b] = artfHipEventCreateOrder()) != NULL;
        for (int k = 0; k < READERS; ++k) ready = ready && (f->reads[b][k] = artfHipEventCreateOrder()) != NULL;
    }
    if (ready) {
        float* p = (float*)f->block;
        f->f.density = p, f->f.radial = p + CELLS, f->f.around = p + 2 * CELLS;
        for (int k = 0; k < 3; ++k) f->f.scratch[k] = p + (3 + k) * CELLS;
        f->fields[0] = p + 6 * CELLS, f->fields[1] = p + 7 * CELLS;
        f->current = 1; /* the first field goes to buffer 0 */
        f->f.field = f->fields[0];
        f->f.rings = (const ArtfFluidRings*)(p + 8 * CELLS);
        ready = !artfHipUpload((void*)f->f.rings, &f->rings, sizeof f->rings, f->stream) && !artfHipFluidSeed(&f->f, f->stream) &&
                eddyRoom(f, FIRST_STIRS);
    }
    if (ready) {
        /* the first stirs, all large, in a row (fluid.c:306), then the warm-up steps at the given time */
        for (int e = 0; e < FIRST_STIRS; ++e) artfFluidDrawEddy(&f->rng, &f->rings, inner, outer, false, &f->eddiesHost[e]);
        ready = !artfHipUpload(f->eddies, f->eddiesHost, FIRST_STIRS * sizeof(ArtfFluidEddy), f->stream) &&
                !artfHipFluidStir(&f->f, f->eddies, FIRST_STIRS, f->stream) && run(f, ARTF_FLUID_WARM);
        f->time = time;
        ready = ready && finish(f, true);
    }
    if (!ready) {
        artfLog("artful: the fluid could not be set up on the device");
        artfDeviceFluidFree(f);
        return NULL;
    }
    return f;
}

static bool advance(ArtfDeviceFluid* f, double time, bool wait) {
    if (time - f->time < ARTF_FLUID_DT) {
        /* gas does not run back; far behind, the clock is simply reset (as the host fluid) */
        if (time < f->time - 8.0) f->time = time;
        return false;
    }
    int steps = (int)((time - f->time) / ARTF_FLUID_DT);
    if (steps > ARTF_FLUID_CATCH_UP) {
        f->time = time - (double)ARTF_FLUID_CATCH_UP * ARTF_FLUID_DT;
        steps = ARTF_FLUID_CATCH_UP;
    }
    if (artfHipSetDevice(f->device) || !run(f, steps) || !finish(f, wait)) {
        artfLog("artful: the fluid's step failed on the device");
        return false;
    }
    return true;
}

bool artfDeviceFluidAdvance(ArtfDeviceFluid* f, double time) { return advance(f, time, true); }

bool artfDeviceFluidAdvanceQueued(ArtfDeviceFluid* f, double time) { return advance(f, time, false); }

const float* artfDeviceFluidField(const ArtfDeviceFluid* f) { return f->fields[f->current]; }

bool artfDeviceFluidReadOn(ArtfDeviceFluid* f, void* stream) { return !artfHipStreamWait(stream, f->written[f->current]); }

bool artfDeviceFluidReadDone(ArtfDeviceFluid* f, void* stream) {
    const int b = f->current;
    if (f->readCount[b] == READERS) { /* no slot left: the fluid's stream waits for the oldest read now */
        if (artfHipStreamWait(f->stream, f->reads[b][0])) return false;
        for (int k = 1; k < READERS; ++k) {
            void* e = f->reads[b][k - 1];
            f->reads[b][k - 1] = f->reads[b][k];
            f->reads[b][k] = e;
        }
        --f->readCount[b];
    }
    return !artfHipEventRecord(f->reads[b][f->readCount[b]++], stream);
}

bool artfDeviceFluidRead(const ArtfDeviceFluid* f, float* field) {
    return !artfHipDownload(field, f->fields[f->current], CELLS * sizeof(float), f->stream);
}

double artfDeviceFluidTime(const ArtfDeviceFluid* f) { return f->time; }
