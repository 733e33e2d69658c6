/* artful-blackhole — the black hole flown around on Linux, without a window: what Artful 0.4.0's
 * tools/blackhole.c does (the eye in orbit at the speed of the gas, seeing what one moving that fast
 * sees; time and the eye's place advanced with the clock; one artfSceneSetBlackHole and one
 * artfRendererDraw a frame, tools/blackhole.c:112-207), with the frames kept in memory, a few of them
 * written out as PPM images, and the timing reported as the reference reports it
 * (artful-blackhole.txt, tools/blackhole.c:196-203).
 *
 *   artful-blackhole [--size w h] [--seconds n | --frames n] [--step dt] [--spin a] [--samples s]
 *                    [--disc-on-device] [--no-orbit] [--doppler d] [--save k dir] [--report file]
 *                    [--describe file]
 *
 * --step dt advances the hole's time by dt a frame (0.2 is one fluid step) instead of by the clock,
 * so that a run is the same frames every time. --disc-on-device keeps the disc's map and fluid on the
 * GPU (artfSceneSetDiscOnDevice). --save k dir writes every k-th frame as dir/frame-NNNNN.ppm, and
 * --describe file a line for each of them with the hole's description and the camera that made it. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artful/device_disc.h"
#include "artful/renderer.h"
#include "artful/scene.h"
#include "frames.h"

typedef struct Flight {
    ArtfScene* scene;
    ArtfBlackHoleDesc hole;
    ArtfCamera camera;
    float around, tilt, distance;
    bool orbiting;
} Flight;

/* The eye on its circle, moving with the gas when in orbit, aimed where the hole appears to it
 * (tools/blackhole.c:35-44). */
static void place(Flight* f) {
    const float c = cosf(f->tilt);
    const ArtfVec3 side = artfVec3(cosf(f->around), 0, -sinf(f->around));
    const float beta = f->orbiting ? artfBlackHoleOrbitSpeed(f->hole.spin, f->distance) : 0.f;
    f->hole.velocity = artfVec3Scale(side, beta);
    const ArtfVec3 aim = artfVec3Scale(f->hole.velocity, f->distance / sqrtf(1.f - beta * beta));
    const ArtfCamera camera = {artfVec3(sinf(f->around) * c * f->distance, sinf(f->tilt) * f->distance, cosf(f->around) * c * f->distance), aim,
                               artfVec3(0, 1, 0), .33f, .1f};
    f->camera = camera;
    artfSceneSetCamera(f->scene, &camera);
}

static void usage(void) {
    fprintf(stderr, "artful-blackhole [--size w h] [--seconds n | --frames n] [--step dt] [--spin a] [--samples s]\n"
                    "                 [--disc-on-device] [--no-orbit] [--doppler d] [--save k dir] [--report file]\n"
                    "                 [--describe file]\n");
}

int main(int argc, char** argv) {
    int width = 1280, height = 720, frames = 0, saveEvery = 0, samples = 1;
    double seconds = 0, step = 0;
    float spin = 0.f, doppler = -1.f;
    bool onDevice = false, orbit = true, spinGiven = false;
    const char* saveDir = NULL;
    const char* report = "artful-blackhole.txt";
    const char* describe = NULL;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!strcmp(a, "--size") && i + 2 < argc) width = atoi(argv[++i]), height = atoi(argv[++i]);
        else if (!strcmp(a, "--seconds") && i + 1 < argc) seconds = atof(argv[++i]);
        else if (!strcmp(a, "--frames") && i + 1 < argc) frames = atoi(argv[++i]);
        else if (!strcmp(a, "--step") && i + 1 < argc) step = atof(argv[++i]);
        else if (!strcmp(a, "--spin") && i + 1 < argc) spin = (float)atof(argv[++i]), spinGiven = true;
        else if (!strcmp(a, "--samples") && i + 1 < argc) samples = atoi(argv[++i]);
        else if (!strcmp(a, "--doppler") && i + 1 < argc) doppler = (float)atof(argv[++i]);
        else if (!strcmp(a, "--disc-on-device")) onDevice = true;
        else if (!strcmp(a, "--no-orbit")) orbit = false;
        else if (!strcmp(a, "--save") && i + 2 < argc) saveEvery = atoi(argv[++i]), saveDir = argv[++i];
        else if (!strcmp(a, "--report") && i + 1 < argc) report = argv[++i];
        else if (!strcmp(a, "--describe") && i + 1 < argc) describe = argv[++i];
        else {
            usage();
            return 2;
        }
    }
    if (width < 1 || height < 1) return usage(), 2;
    if (seconds <= 0 && frames <= 0) frames = 100;

    ArtfRendererOptions o = artfRendererDefaults(width, height);
    o.backend = ARTF_BACKEND_HIP;
    ArtfRenderer* renderer = artfRendererNew(&o);
    if (!renderer) {
        fprintf(stderr, "artful-blackhole: no HIP device\n");
        return 1;
    }
    Flight f;
    memset(&f, 0, sizeof f);
    f.scene = artfSceneNew();
    if (onDevice) artfSceneSetDiscOnDevice(f.scene, 0);
    f.hole = artfBlackHoleDefaults();
    if (spinGiven) f.hole.spin = spin;
    if (doppler >= 0) f.hole.doppler = doppler;
    f.hole.samples = samples;
    artfSceneSetBlackHole(f.scene, &f.hole);
    f.tilt = .06f, f.distance = 74.1f, f.orbiting = orbit;

    uint32_t* pixels = malloc((size_t)width * height * sizeof(uint32_t));
    FILE* described = describe ? fopen(describe, "w") : NULL;
    const double start = artfNowMs();
    double last = start, gpuMs = 0;
    int drawn = 0;
    for (;;) {
        const double now = artfNowMs();
        const float passed = step > 0 ? (float)step : (float)(now - last) * .012f;
        f.hole.time += passed;
        /* the eye's clock and the gas's are the same, so it keeps the pace an orbit has */
        if (f.orbiting) f.around += passed / (f.distance * sqrtf(f.distance) + f.hole.spin);
        last = now;
        place(&f);
        artfSceneSetBlackHole(f.scene, &f.hole);
        if (!artfRendererDraw(renderer, f.scene)) {
            fprintf(stderr, "artful-blackhole: the device was lost\n");
            break;
        }
        ArtfHoleStats st;
        if (artfRendererHoleStats(renderer, &st)) gpuMs += st.traceMs + st.resolveMs + st.bloomMs + st.toneMs;
        if (saveDir && saveEvery > 0 && drawn % saveEvery == 0) {
            char path[1024];
            snprintf(path, sizeof path, "%s/frame-%05d.ppm", saveDir, drawn);
            artfRendererReadPixels(renderer, pixels);
            if (!driverSaveImage(path, pixels, width, height)) fprintf(stderr, "artful-blackhole: could not write into %s\n", saveDir);
            driverDescribe(described, drawn, &f.hole, &f.camera);
        }
        ++drawn;
        if (frames > 0 && drawn >= frames) break;
        if (seconds > 0 && artfNowMs() - start > seconds * 1000) break;
    }
    const double wall = artfNowMs() - start;
    FILE* out = fopen(report, "w");
    if (out) {
        fprintf(out, "%s, %d x %d: %d frames, %.2f ms a frame (GPU %.2f ms), disc on the %s\n", artfRendererDescription(renderer), width, height, drawn,
                wall / (drawn ? drawn : 1), gpuMs / (drawn ? drawn : 1), onDevice ? "device" : "host");
        fclose(out);
    }
    printf("%d frames, %.2f ms a frame\n", drawn, wall / (drawn ? drawn : 1));
    if (described) fclose(described);
    free(pixels);
    artfSceneFree(f.scene);
    artfRendererFree(renderer);
    return 0;
}
