/* The host code of the library (memory, pool, disc map, fluid, frame, scene) built on its own, for the
 * sanitizer and determinism test (tests/test_hygiene.py): gcc with -fsanitize=address,undefined and
 * ARTFUL_COUNT_ALLOCATIONS, no HIP. For each thread count given, a scene is made, its hole set at each
 * time given (the first makes the map and fluid, the others move the fluid on), its map, fluid and
 * frame written to files, and the scene freed; at the end the blocks still alive are printed, which
 * must be none.
 *
 *   runner <out dir> <desc hex> <camera hex> <width> <height> <times, comma-separated> <threads>...
 *
 * This build has no GPU: the calls that would keep a disc on one report that no device can hold it,
 * and a scene never asks for one here (artfSceneSetDiscOnDevice is not called). */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artful/device_disc.h"
#include "core/pool.h"
#include "scene/scene_internal.h"

long artfAllocationsAlive(void);

float* artfDeviceMapNew(int device) { (void)device; return NULL; }
bool artfDeviceMapRings(float* map, float inner, float outer, float spin, int device) { (void)map, (void)inner, (void)outer, (void)spin, (void)device; return false; }
void artfDeviceMapFree(float* map) { (void)map; }
ArtfDeviceFluid* artfDeviceFluidNew(float inner, float outer, float spin, double time, int device) { (void)inner, (void)outer, (void)spin, (void)time, (void)device; return NULL; }
void artfDeviceFluidFree(ArtfDeviceFluid* f) { (void)f; }
bool artfDeviceFluidAdvanceQueued(ArtfDeviceFluid* f, double time) { (void)f, (void)time; return false; }
double artfDeviceFluidTime(const ArtfDeviceFluid* f) { (void)f; return 0; }

static void unhex(const char* text, void* out, size_t n) {
    unsigned char* b = out;
    if (strlen(text) != 2 * n) {
        fprintf(stderr, "runner: %zu bytes of hex expected\n", n);
        exit(2);
    }
    for (size_t i = 0; i < n; ++i) {
        unsigned v = 0;
        sscanf(text + 2 * i, "%2x", &v);
        b[i] = (unsigned char)v;
    }
}

static void put(const char* dir, int threads, const char* what, const void* p, size_t bytes) {
    char path[1024];
    snprintf(path, sizeof path, "%s/%s-%d.bin", dir, what, threads);
    FILE* f = fopen(path, "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes || fclose(f)) {
        fprintf(stderr, "runner: cannot write %s\n", path);
        exit(1);
    }
}

int main(int argc, char** argv) {
    if (argc < 8) {
        fputs("runner <out dir> <desc hex> <camera hex> <width> <height> <times> <threads>...\n", stderr);
        return 2;
    }
    ArtfBlackHoleDesc desc;
    ArtfCamera camera;
    unhex(argv[2], &desc, sizeof desc);
    unhex(argv[3], &camera, sizeof camera);
    const int width = atoi(argv[4]), height = atoi(argv[5]);
    for (int a = 7; a < argc; ++a) {
        const int threads = atoi(argv[a]);
        artfPoolSetDefaultSize(threads);
        ArtfScene* s = artfSceneNew();
        artfSceneSetCamera(s, &camera);
        char* times = strdup(argv[6]);
        for (char* t = strtok(times, ","); t; t = strtok(NULL, ",")) {
            desc.time = (float)atof(t);
            artfSceneSetBlackHole(s, &desc);
        }
        free(times);
        ArtfHoleFrame frame;
        uint32_t bloomPixels = 0;
        artfHoleFrame(s, width, height, &frame, &bloomPixels);
        put(argv[1], threads, "map", s->hole.map, (size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float));
        put(argv[1], threads, "fluid", artfFluidField(s->hole.fluid), (size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT * sizeof(float));
        put(argv[1], threads, "frame", &frame, sizeof frame);
        artfSceneFree(s);
    }
    artfPoolSetDefaultSize(0);
    printf("allocations alive %ld\n", artfAllocationsAlive());
    return 0;
}
