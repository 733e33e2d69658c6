/* artful-hip — the black hole's description, its clamps, and the uniforms of a frame.
 *
 * artfBlackHoleDefaults, artfSceneSetBlackHole, artfHoleIsco, artfBlackHoleOrbitSpeed and
 * artfHoleFrame keep the behaviour of Artful 0.4.0's src/scene/hole.c:200-371: the same defaults, the
 * same clamps, the same rebuild rules, and a frame whose 368 bytes are the reference's bit for bit (its
 * arithmetic order is kept for that; tanf is the C library's, as there). artfHoleFrame64 works the same
 * geometry out in doubles (frame64.c), for the FP64 and mixed-precision tracers. */
#include "scene_internal.h"

#include "artful/device_disc.h"
#include "frame_body.h"

static double cubeRootNewton(double x) {
    if (x <= 0) return 0;
    double v = x > 1 ? x : 1;
    for (int i = 0; i < 60; ++i) v = (2.0 * v + x / (v * v)) / 3.0;
    return v;
}

/* Bardeen, Press and Teukolsky (1972), eq. 2.21: the last stable circular orbit. */
float artfHoleIsco(float spin) {
    const double a = spin, a2 = a * a;
    const double z1 = 1.0 + cubeRootNewton(1.0 - a2) * (cubeRootNewton(1.0 + a) + cubeRootNewton(1.0 - a));
    const double z2 = sqrt(3.0 * a2 + z1 * z1);
    const double root = sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
    return (float)(3.0 + z2 - (a < 0 ? -root : root));
}

/* Orbital angular velocity 1/(r^1.5 + a) against the frame drag 2ar/A, times varpi over the lapse. */
float artfBlackHoleOrbitSpeed(float spin, float r) {
    const double a = spin, rr = (double)r * r, delta = rr - 2.0 * r + a * a;
    const double bigA = (rr + a * a) * (rr + a * a) - a * a * delta;
    const double omega = 1.0 / ((double)r * sqrt((double)r) + a), drag = 2.0 * a * r / bigA;
    const double varpi = sqrt(bigA) / r, lapse = r * sqrt(delta) / sqrt(bigA);
    const double v = (omega - drag) * varpi / lapse;
    return (float)(v > .99 ? .99 : v);
}

ArtfBlackHoleDesc artfBlackHoleDefaults(void) {
    ArtfBlackHoleDesc d;
    memset(&d, 0, sizeof d);
    d.axis = artfVec3(0, -1, 0);
    d.mass = 1.f;
    d.spin = .6f;
    d.discOuter = 34.f;
    d.discHeight = .03f;
    d.discThickness = 1.f;
    d.discGain = 2.1f;
    d.stars = .4f;
    d.bloom = 4.5f;
    d.exposure = 1.f;
    d.samples = 1;
    return d;
}

/* The description as a scene keeps it (hole.c:245-258): spin within +-.998, the inner edge the last
 * stable orbit or at least .2 past the horizon, the outer at least a mass further, samples 1..8. */
ArtfBlackHoleDesc artfHoleClamped(const ArtfBlackHoleDesc* desc) {
    ArtfBlackHoleDesc d = *desc;
    d.spin = artfClampf(desc->spin, -.998f, .998f);
    const float horizon = 1.f + sqrtf(1.f - d.spin * d.spin);
    d.discInner = desc->discInner > 0 ? artfMaxf(desc->discInner, horizon + .2f) : artfHoleIsco(d.spin);
    d.discOuter = artfMaxf(desc->discOuter, d.discInner + 1.f);
    d.mass = artfMaxf(desc->mass, 1e-6f);
    d.discHeight = artfMaxf(desc->discHeight, 1e-3f);
    d.samples = artfClampi(desc->samples, 1, 8);
    return d;
}

/* A new fluid for the hole's disc at this moment: on the host, or on the scene's disc device (falling
 * back to the host, logged, if that device cannot hold it). */
static void newFluid(ArtfHole* h, double time) {
    const float inner = h->desc.discInner, outer = h->desc.discOuter, spin = h->desc.spin;
    artfFluidFree(h->fluid);
    artfDeviceFluidFree(h->deviceFluid);
    h->fluid = NULL, h->deviceFluid = NULL;
    if (h->discDevice >= 0 && !(h->deviceFluid = artfDeviceFluidNew(inner, outer, spin, time, h->discDevice)))
        artfLog("artful: the fluid stays on the host");
    if (!h->deviceFluid) h->fluid = artfFluidNew(inner, outer, spin, time);
    ++h->fluidVersion;
}

/* The map for the hole's disc: painted once (on the disc device, or on the host as the reference
 * does, also when that device cannot hold it), then this disc's ring tables. */
static void ringMap(ArtfHole* h) {
    const float inner = h->desc.discInner, outer = h->desc.discOuter, spin = h->desc.spin;
    if (h->discDevice >= 0 && !h->deviceMap && !h->map && !(h->deviceMap = artfDeviceMapNew(h->discDevice)))
        artfLog("artful: the disc map stays on the host");
    if (h->deviceMap && !artfDeviceMapRings(h->deviceMap, inner, outer, spin, h->discDevice)) {
        artfLog("artful: the disc map goes back to the host");
        artfDeviceMapFree(h->deviceMap);
        h->deviceMap = NULL;
    }
    if (!h->deviceMap) {
        if (!h->map) {
            h->map = artfMalloc((size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float));
            artfDiscPaint(h->map);
        }
        artfDiscRings(h->map, inner, outer, spin);
    }
    ++h->mapVersion;
}

void artfSceneSetBlackHole(ArtfScene* s, const ArtfBlackHoleDesc* desc) {
    ArtfHole* h = &s->hole;
    if (!desc) {
        h->on = false;
        return;
    }
    const ArtfBlackHoleDesc clamped = artfHoleClamped(desc);
    const float spin = clamped.spin, inner = clamped.discInner, outer = clamped.discOuter;
    const bool sameDisc = (h->map || h->deviceMap) && h->desc.discInner == inner && h->desc.discOuter == outer;
    const bool sameSpin = sameDisc && h->desc.spin == spin;
    h->on = true;
    h->desc = clamped;
    if (!sameSpin) {
        ringMap(h);
        newFluid(h, desc->time);
    } else if (h->deviceFluid ? artfDeviceFluidAdvanceQueued(h->deviceFluid, desc->time) : artfFluidAdvance(h->fluid, desc->time)) {
        ++h->fluidVersion;
    }
}

double artfHoleFluidTime(const ArtfHole* h) {
    if (h->deviceFluid) return artfDeviceFluidTime(h->deviceFluid);
    return h->fluid ? artfFluidTime(h->fluid) : 0.0;
}

void artfSceneSetDiscOnDevice(ArtfScene* s, int device) {
    ArtfHole* h = &s->hole;
    if (device < 0) device = -1;
    if (device == h->discDevice) return;
    const bool had = h->map || h->deviceMap;
    h->discDevice = device;
    artfFree(h->map);
    artfDeviceMapFree(h->deviceMap);
    h->map = NULL, h->deviceMap = NULL;
    if (had) { /* the disc again, in its new place; its gas starts again */
        ringMap(h);
        newFluid(h, h->desc.time);
    }
}

void artfHoleFrameOf(const ArtfBlackHoleDesc* d, const ArtfCamera* c, double fluidTime, int width, int height, int levelsCap, ArtfHoleFrame* out,
                     uint32_t* bloomPixels) {
    memset(out, 0, sizeof *out);
#define CAMERA_EYE(k) c->position.k
    HOLE_FRAME_BODY(float, sqrtf, tanf, fabsf, d->spin, CAMERA_EYE)
    out->flow[0] = (float)fluidTime;
    out->look[0] = d->discGain, out->look[1] = d->doppler, out->look[2] = d->stars, out->look[3] = d->bloom;
    out->image[0] = d->exposure;
    out->image[1] = 2.f / (float)width;
    out->image[2] = 2.f / (float)height;
    out->image[3] = 1.f / (float)(d->samples * d->samples);
    out->size[0] = (uint32_t)width, out->size[1] = (uint32_t)height, out->size[2] = (uint32_t)d->samples;
    /* Bloom levels, each half the last, while both sides are above 3; always at least one, since the
     * tone map reads it (hole.c:358-370). */
    const uint32_t most = levelsCap > 0 && levelsCap < ARTF_BLOOM_LEVELS ? (uint32_t)levelsCap : ARTF_BLOOM_LEVELS;
    uint32_t lw = (uint32_t)width, lh = (uint32_t)height, at = 0, levels = 0;
    while (levels < most && (!levels || (lw > 3 && lh > 3))) {
        lw = (lw + 1) / 2;
        lh = (lh + 1) / 2;
        out->bloom[levels][0] = at, out->bloom[levels][1] = lw, out->bloom[levels][2] = lh;
        at += lw * lh;
        ++levels;
    }
    out->size[3] = levels;
    if (bloomPixels) *bloomPixels = at;
}

void artfHoleFrameLevels(const ArtfScene* s, int width, int height, int levelsCap, ArtfHoleFrame* out, uint32_t* bloomPixels) {
    const double t = artfHoleFluidTime(&s->hole);
    artfHoleFrameOf(&s->hole.desc, &s->camera, t, width, height, levelsCap, out, bloomPixels);
}

void artfHoleFrame(const ArtfScene* s, int width, int height, ArtfHoleFrame* out, uint32_t* bloomPixels) {
    artfHoleFrameLevels(s, width, height, 0, out, bloomPixels);
}

void artfHoleFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out) { artfHoleFrame64Of(&s->hole.desc, &s->camera, width, height, out); }
