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
    return !artfHipSetDevice(device) && !artfHipUpload(map + ARTF_HOLE_ROWS_AT, rows, sizeof rows, NULL) &&
           !artfHipUpload(map + ARTF_HOLE_GLOW_AT, glow, sizeof glow, NULL) && !artfHipStreamSync(NULL);
}

void artfDeviceMapFree(float* map) { artfHipFree(map); }

struct ArtfDeviceFluid {
    int device;
    void* stream;
    float inner, outer;
    ArtfFluidRings rings;
    uint32_t rng;
    double time;
    void* block; /* device: the seven fields, then the ring constants */
    ArtfHipFluid f;
    ArtfFluidEddy* eddies; /* device, room for eddyRoom */
    int eddyRoom;
    ArtfFluidEddy* eddiesHost;
};

void artfDeviceFluidFree(ArtfDeviceFluid* f) {
    if (!f) return;
    if (f->stream) artfHipStreamSync(f->stream);
    artfHipFree(f->block);
    artfHipFree(f->eddies);
    artfFree(f->eddiesHost);
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

/* steps steps: their eddies drawn in order on the host, sent at once, then the passes. */
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

static bool finish(ArtfDeviceFluid* f) { return !artfHipFluidField(&f->f, f->stream) && !artfHipStreamSync(f->stream); }

ArtfDeviceFluid* artfDeviceFluidNew(float inner, float outer, float spin, double time, int device) {
    if (device < 0 || device >= artfHipDeviceCount() || artfHipSetDevice(device)) {
        artfLog("artful: no HIP device %d for the fluid", device);
        return NULL;
    }
    ArtfDeviceFluid* f = ARTF_NEW(ArtfDeviceFluid);
    f->device = device, f->inner = inner, f->outer = outer, f->rng = ARTF_FLUID_SEED, f->time = time;
    artfFluidRingsOf(inner, outer, spin, &f->rings);
    f->stream = artfHipStreamCreate();
    f->block = artfHipMalloc(7 * CELLS * sizeof(float) + sizeof(ArtfFluidRings));
    bool ready = f->stream && f->block;
    if (ready) {
        float* p = (float*)f->block;
        f->f.density = p, f->f.radial = p + CELLS, f->f.around = p + 2 * CELLS;
        for (int k = 0; k < 3; ++k) f->f.scratch[k] = p + (3 + k) * CELLS;
        f->f.field = p + 6 * CELLS;
        f->f.rings = (const ArtfFluidRings*)(p + 7 * CELLS);
        ready = !artfHipUpload((void*)f->f.rings, &f->rings, sizeof f->rings, f->stream) && !artfHipFluidSeed(&f->f, f->stream) &&
                eddyRoom(f, FIRST_STIRS);
    }
    if (ready) {
        /* the first stirs, all large, in a row (fluid.c:306), then the warm-up steps at the given time */
        for (int e = 0; e < FIRST_STIRS; ++e) artfFluidDrawEddy(&f->rng, &f->rings, inner, outer, false, &f->eddiesHost[e]);
        ready = !artfHipUpload(f->eddies, f->eddiesHost, FIRST_STIRS * sizeof(ArtfFluidEddy), f->stream) &&
                !artfHipFluidStir(&f->f, f->eddies, FIRST_STIRS, f->stream) && run(f, ARTF_FLUID_WARM);
        f->time = time;
        ready = ready && finish(f);
    }
    if (!ready) {
        artfLog("artful: the fluid could not be set up on the device");
        artfDeviceFluidFree(f);
        return NULL;
    }
    return f;
}

bool artfDeviceFluidAdvance(ArtfDeviceFluid* f, double time) {
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
    if (artfHipSetDevice(f->device) || !run(f, steps) || !finish(f)) {
        artfLog("artful: the fluid's step failed on the device");
        return false;
    }
    return true;
}

const float* artfDeviceFluidField(const ArtfDeviceFluid* f) { return f->f.field; }

bool artfDeviceFluidRead(const ArtfDeviceFluid* f, float* field) { return !artfHipDownload(field, f->f.field, CELLS * sizeof(float), f->stream); }

double artfDeviceFluidTime(const ArtfDeviceFluid* f) { return f->time; }
