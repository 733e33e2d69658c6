/* artful-render — Artful 0.4.0's headless tool (tools/render.c) for its black-hole scene, through the HIP
 * backend: the scene, its defaults and options as there (render.c:333-360, 398, 435-452), one picture
 * or a sequence of them, and with --frames the reference's timing (render.c:542-580): the hole's time
 * goes on by .2 a frame, so each frame also takes one fluid step, and a GPU's frames are timed as a run
 * to the last pixel, with reading the pixels back taken out.
 *
 *   artful-render [--backend hip|auto] [--size w h] [--scene blackhole] [--frames n] [--out file.png|.ppm]
 *                 [--time t] [--samples n] [--doppler 0..1] [--distance masses] [--spin a] [--tilt degrees]
 *                 [--roll degrees] [--fov degrees] [--exposure x] [--bloom x] [--around degrees]
 *                 [--shutter t] [--orbit] [--sequence n --step t --out frames/%04d.png] [--to-distance d]
 *                 [--to-tilt t] [--to-around a] [--precision fp32|fp32-fast|mixed|fp64] [--describe file]
 *                 [--disc-on-device]
 *
 * Only the black-hole scene is here: the reference's triangle scenes are not on this backend's path.
 * Beyond the reference: --precision picks the tracer; --describe writes, for each picture written, the
 * hole's description and the camera that made it; --disc-on-device keeps the disc's map and fluid on the
 * GPU (artfSceneSetDiscOnDevice: the same pictures, the fluid's step a frame taken off the CPU). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artful/device_disc.h"
#include "artful/renderer.h"
#include "artful/scene.h"
#include "frames.h"

/* the black hole's options, with the reference's defaults: the eye of the film's own pictures */
typedef struct HoleOptions {
    float time, doppler, distance, tilt, roll, fov, exposure, spin, around, shutter, bloom;
    int samples;
    bool orbit;
    int sequence;
    float step, toDistance, toTilt, toAround;
} HoleOptions;

typedef struct Made {
    ArtfBlackHoleDesc desc;
    ArtfCamera camera;
} Made;

/* The scene: the hole with the given options, the eye `distance` away at `tilt` above the disc and
 * `around` about its axis, rolled by `roll`; moving at the gas's orbital speed with --orbit, and then
 * aimed where the hole appears to one moving so (render.c:336-360). */
static Made sceneBlackHole(ArtfScene* s, const HoleOptions* h) {
    Made m;
    ArtfBlackHoleDesc d = artfBlackHoleDefaults();
    d.time = h->time;
    d.doppler = h->doppler;
    d.samples = h->samples;
    d.exposure = h->exposure;
    d.spin = h->spin;
    d.shutter = h->shutter;
    if (h->bloom >= 0) d.bloom = h->bloom;
    const float tilt = h->tilt * 3.14159265f / 180.f, roll = h->roll * 3.14159265f / 180.f, around = h->around * 3.14159265f / 180.f;
    const float flat = cosf(tilt) * h->distance;
    if (h->orbit) {
        const float speed = artfBlackHoleOrbitSpeed(h->spin, h->distance);
        d.velocity = artfVec3(cosf(around) * speed, 0, -sinf(around) * speed);
    }
    artfSceneSetBlackHole(s, &d);
    const ArtfVec3 eye = artfVec3(sinf(around) * flat, sinf(tilt) * h->distance, cosf(around) * flat);
    const ArtfVec3 side = artfVec3(cosf(around), 0, -sinf(around));
    const ArtfVec3 up = artfVec3Add(artfVec3Scale(side, sinf(roll)), artfVec3(0, cosf(roll), 0));
    const float beta = artfVec3Length(d.velocity);
    const ArtfVec3 aim = artfVec3Scale(d.velocity, h->distance / sqrtf(1.f - beta * beta));
    const ArtfCamera c = {eye, aim, up, h->fov * 3.14159265f / 180.f, .1f};
    artfSceneSetCamera(s, &c);
    m.desc = d;
    m.camera = c;
    return m;
}

static void printLog(void* user, const char* message) {
    (void)user;
    fprintf(stderr, "%s\n", message);
}

static void usage(void) {
    fputs("artful-render [--backend hip|auto] [--size w h] [--scene blackhole] [--frames n] [--out file.png|.ppm]\n"
          "              [--time t] [--samples n] [--doppler d] [--distance r] [--spin a] [--tilt deg] [--roll deg] [--fov deg]\n"
          "              [--exposure x] [--bloom x] [--around deg] [--shutter t] [--orbit] [--sequence n --step t]\n"
          "              [--to-distance d] [--to-tilt t] [--to-around a] [--precision fp32|fp32-fast|mixed|fp64] [--describe file]\n"
          "              [--disc-on-device]\n",
          stderr);
}

int main(int argc, char** argv) {
    artfSetLog(printLog, NULL);
    ArtfRendererOptions o = artfRendererDefaults(1280, 720);
    o.backend = ARTF_BACKEND_HIP;
    o.noVsync = true;
    const char *out = NULL, *describe = NULL;
    int frames = 0;
    bool onDevice = false;
    ArtfHoleTracing tracing = artfHoleTracingDefaults();
    HoleOptions hole = {0, 0, 74.1f, 3.44f, 0, 19.f, 1.f, .6f, 0, 0, -1.f, 1, false, 0, .9f, -1.f, -1000.f, -1000.f};
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!strcmp(a, "--backend") && i + 1 < argc) {
            ++i;
            if (!strcmp(argv[i], "hip")) o.backend = ARTF_BACKEND_HIP;
            else if (!strcmp(argv[i], "auto")) o.backend = ARTF_BACKEND_AUTO;
            else {
                fprintf(stderr, "artful-render: the %s backend is not in this library (hip, auto)\n", argv[i]);
                return 1;
            }
        } else if (!strcmp(a, "--size") && i + 2 < argc) {
            o.width = atoi(argv[++i]);
            o.height = atoi(argv[++i]);
        } else if (!strcmp(a, "--scene") && i + 1 < argc) {
            if (strcmp(argv[++i], "blackhole")) {
                fprintf(stderr, "artful-render: only the blackhole scene is on this backend's path\n");
                return 2;
            }
        } else if (!strcmp(a, "--frames") && i + 1 < argc) frames = atoi(argv[++i]);
        else if (!strcmp(a, "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(a, "--describe") && i + 1 < argc) describe = argv[++i];
        else if (!strcmp(a, "--time") && i + 1 < argc) hole.time = (float)atof(argv[++i]);
        else if (!strcmp(a, "--samples") && i + 1 < argc) hole.samples = atoi(argv[++i]);
        else if (!strcmp(a, "--doppler") && i + 1 < argc) hole.doppler = (float)atof(argv[++i]);
        else if (!strcmp(a, "--distance") && i + 1 < argc) hole.distance = (float)atof(argv[++i]);
        else if (!strcmp(a, "--tilt") && i + 1 < argc) hole.tilt = (float)atof(argv[++i]);
        else if (!strcmp(a, "--roll") && i + 1 < argc) hole.roll = (float)atof(argv[++i]);
        else if (!strcmp(a, "--fov") && i + 1 < argc) hole.fov = (float)atof(argv[++i]);
        else if (!strcmp(a, "--exposure") && i + 1 < argc) hole.exposure = (float)atof(argv[++i]);
        else if (!strcmp(a, "--spin") && i + 1 < argc) hole.spin = (float)atof(argv[++i]);
        else if (!strcmp(a, "--around") && i + 1 < argc) hole.around = (float)atof(argv[++i]);
        else if (!strcmp(a, "--shutter") && i + 1 < argc) hole.shutter = (float)atof(argv[++i]);
        else if (!strcmp(a, "--orbit")) hole.orbit = true;
        else if (!strcmp(a, "--disc-on-device")) onDevice = true;
        else if (!strcmp(a, "--bloom") && i + 1 < argc) hole.bloom = (float)atof(argv[++i]);
        else if (!strcmp(a, "--sequence") && i + 1 < argc) hole.sequence = atoi(argv[++i]);
        else if (!strcmp(a, "--step") && i + 1 < argc) hole.step = (float)atof(argv[++i]);
        else if (!strcmp(a, "--to-distance") && i + 1 < argc) hole.toDistance = (float)atof(argv[++i]);
        else if (!strcmp(a, "--to-tilt") && i + 1 < argc) hole.toTilt = (float)atof(argv[++i]);
        else if (!strcmp(a, "--to-around") && i + 1 < argc) hole.toAround = (float)atof(argv[++i]);
        else if (!strcmp(a, "--precision") && i + 1 < argc) {
            const char* p = argv[++i];
            tracing.precision = !strcmp(p, "fp64") ? ARTF_PRECISION_FP64 : !strcmp(p, "mixed") ? ARTF_PRECISION_MIXED
                              : !strcmp(p, "fp32-fast") ? ARTF_PRECISION_FP32_FAST : ARTF_PRECISION_FP32;
        } else {
            fprintf(stderr, "unknown: %s\n", a);
            usage();
            return 2;
        }
    }
    if (o.width < 1 || o.height < 1) return usage(), 2;

    ArtfRenderer* r = artfRendererNew(&o);
    if (!r) {
        fputs("that backend cannot run here\n", stderr);
        return 1;
    }
    artfRendererSetHoleTracing(r, &tracing);
    printf("%s, %d x %d\n", artfRendererDescription(r), o.width, o.height);
    FILE* described = describe ? fopen(describe, "w") : NULL;
    uint32_t* pixels = malloc((size_t)o.width * o.height * 4);
    ArtfScene* s = artfSceneNew();
    if (onDevice) artfSceneSetDiscOnDevice(s, 0);
    const double t0 = artfNowMs();
    Made made = sceneBlackHole(s, &hole);
    printf("scene built in %.1f ms\n", artfNowMs() - t0);

    int status = 0;
    if (hole.sequence > 0 && out) {
        /* n pictures, the gas flowing from one to the next, the eye going smoothly to its last place */
        const HoleOptions from = hole;
        const float toDistance = from.toDistance > 0 ? from.toDistance : from.distance;
        const float toTilt = from.toTilt > -999.f ? from.toTilt : from.tilt, toAround = from.toAround > -999.f ? from.toAround : from.around;
        for (int i = 0; i < hole.sequence && !status; ++i) {
            const float x = hole.sequence > 1 ? (float)i / (float)(hole.sequence - 1) : 0.f, e = x * x * (3.f - 2.f * x);
            hole.distance = from.distance + (toDistance - from.distance) * e;
            hole.tilt = from.tilt + (toTilt - from.tilt) * e;
            hole.around = from.around + (toAround - from.around) * e;
            hole.time = from.time + from.step * (float)i;
            made = sceneBlackHole(s, &hole);
            char path[1024];
            snprintf(path, sizeof path, out, i);
            if (!artfRendererDraw(r, s) || !artfRendererReadPixels(r, pixels) || !driverSaveImage(path, pixels, o.width, o.height)) {
                fputs("could not write the picture\n", stderr);
                status = 1;
            }
            driverDescribe(described, i, &made.desc, &made.camera);
        }
    } else if (!artfRendererDraw(r, s)) {
        fputs("the device was lost\n", stderr);
        status = 1;
    } else {
        if (frames > 0) {
            double lens = 0, raster = 0, best = 1e9;
            ArtfFrameStats st;
            memset(&st, 0, sizeof st);
            /* A GPU's frame ends when its pixels can be read, not when the call returns: the frames are
             * timed as a run, to the last pixel, and reading itself is taken out (render.c:545-575). */
            artfRendererReadPixels(r, pixels);
            const double start = artfNowMs();
            for (int i = 0; i < frames; ++i) {
                hole.time += .2f;
                made = sceneBlackHole(s, &hole);
                const double a = artfNowMs();
                if (!artfRendererDraw(r, s)) {
                    fputs("the device was lost\n", stderr);
                    status = 1;
                    break;
                }
                const double b = artfNowMs() - a;
                best = b < best ? b : best;
            }
            artfRendererReadPixels(r, pixels);
            double total = artfNowMs() - start;
            const double a = artfNowMs();
            artfRendererReadPixels(r, pixels);
            total -= artfNowMs() - a;
            artfRendererStats(r, &st);
            lens = st.lensMs, raster = st.rasterMs;
            printf("%d frames: %.3f ms a frame (%.0f fps), %.3f ms of it on the CPU at best, %.3f ms on the GPU the last frame "
                   "(lens %.3f, tone map %.3f)\n",
                   frames, total / frames, 1000.0 * frames / total, best, st.gpuMs, lens, raster);
        }
        if (out && !status) {
            if (!artfRendererReadPixels(r, pixels) || !driverSaveImage(out, pixels, o.width, o.height)) {
                fputs("could not write the picture\n", stderr);
                status = 1;
            } else {
                driverDescribe(described, 0, &made.desc, &made.camera);
                printf("wrote %s\n", out);
            }
        }
    }
    if (described) fclose(described);
    free(pixels);
    artfRendererFree(r);
    artfSceneFree(s);
    return status;
}
