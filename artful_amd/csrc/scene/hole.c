/* artful-hip — the black hole's description, its clamps, and the uniforms of a frame.
 *
 * artfBlackHoleDefaults, artfSceneSetBlackHole, artfHoleIsco, artfBlackHoleOrbitSpeed and
 * artfHoleFrame keep the behaviour of Artful 0.4.0's src/scene/hole.c:200-371: the same defaults, the
 * same clamps, the same rebuild rules, and a frame whose 368 bytes are the reference's bit for bit (its
 * arithmetic order is kept for that; tanf is the C library's, as there). artfHoleFrame64 works the same
 * geometry out in doubles, for the FP64 and mixed-precision tracers. */
#include "scene_internal.h"

#include "artful/device_disc.h"

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

/* The eye's local frame and the Kerr metric there, in either precision. HOLE_REAL is float or double;
 * the float form is the reference's arithmetic order exactly. */
#define HOLE_FRAME_BODY(HOLE_REAL, SQRT, TAN, FABS, SPIN, EYE)                                               \
    const HOLE_REAL a = SPIN, horizon = (HOLE_REAL)1 + SQRT((HOLE_REAL)1 - a * a);                        \
    /* the hole's own axes: the disc in x-z, its axis y */                                                   \
    HOLE_REAL gx = d->axis.x, gy = d->axis.y, gz = d->axis.z;                                                \
    {                                                                                                        \
        const HOLE_REAL gl = SQRT(gx * gx + gy * gy + gz * gz);                                              \
        if (gl > 0) {                                                                                        \
            const HOLE_REAL inv = (HOLE_REAL)1 / gl;                                                         \
            gx *= inv, gy *= inv, gz *= inv;                                                                 \
        } else {                                                                                             \
            gx = 0, gy = 1, gz = 0;                                                                          \
        }                                                                                                    \
    }                                                                                                        \
    const HOLE_REAL ox = FABS(gz) < (HOLE_REAL).9f ? 0 : 1, oz = FABS(gz) < (HOLE_REAL).9f ? 1 : 0;          \
    HOLE_REAL exx = gy * oz - gz * 0, exy = gz * ox - gx * oz, exz = gx * 0 - gy * ox;                        \
    {                                                                                                        \
        const HOLE_REAL l = SQRT(exx * exx + exy * exy + exz * exz);                                         \
        if (l > 0) {                                                                                         \
            const HOLE_REAL inv = (HOLE_REAL)1 / l;                                                          \
            exx *= inv, exy *= inv, exz *= inv;                                                              \
        }                                                                                                    \
    }                                                                                                        \
    const HOLE_REAL eyx = gx, eyy = gy, eyz = gz;                                                            \
    const HOLE_REAL ezx = exy * eyz - exz * eyy, ezy = exz * eyx - exx * eyz, ezz = exx * eyy - exy * eyx;    \
    const HOLE_REAL invMass = (HOLE_REAL)1 / (HOLE_REAL)d->mass;                                             \
    HOLE_REAL vx = ((HOLE_REAL)EYE(x) - (HOLE_REAL)d->position.x) * invMass;                                 \
    HOLE_REAL vy = ((HOLE_REAL)EYE(y) - (HOLE_REAL)d->position.y) * invMass;                                 \
    HOLE_REAL vz = ((HOLE_REAL)EYE(z) - (HOLE_REAL)d->position.z) * invMass;                                 \
    HOLE_REAL r = SQRT(vx * vx + vy * vy + vz * vz);                                                         \
    const HOLE_REAL least = horizon + (HOLE_REAL).3f;                                                        \
    if (r < least) { /* the eye stays just outside the horizon */                                           \
        if (r > 0) {                                                                                         \
            const HOLE_REAL k = least / r;                                                                   \
            vx *= k, vy *= k, vz *= k;                                                                       \
        } else {                                                                                             \
            vx = 0, vy = 0, vz = least;                                                                      \
        }                                                                                                    \
        r = least;                                                                                           \
    }                                                                                                        \
    const HOLE_REAL px = (vx * exx + vy * exy + vz * exz) / r;                                               \
    const HOLE_REAL py = (vx * eyx + vy * eyy + vz * eyz) / r;                                               \
    const HOLE_REAL pz = (vx * ezx + vy * ezy + vz * ezz) / r;                                               \
    const HOLE_REAL sine = SQRT(px * px + pz * pz);                                                          \
    const HOLE_REAL cx = sine > (HOLE_REAL)1e-5f ? px / sine : 1, cz = sine > (HOLE_REAL)1e-5f ? pz / sine : 0; \
    const HOLE_REAL er[3] = {px, pz, py}, et[3] = {py * cx, py * cz, -sine}, ep[3] = {-cz, cx, 0};           \
    out->er[0] = er[0], out->er[1] = er[1], out->er[2] = er[2], out->er[3] = r;                              \
    out->etheta[0] = et[0], out->etheta[1] = et[1], out->etheta[2] = et[2];                                  \
    out->ephi[0] = ep[0], out->ephi[1] = ep[1], out->ephi[2] = ep[2];                                        \
    /* Kerr at the eye, for one at rest in the space that turns there (a ZAMO) */                            \
    const HOLE_REAL mu = py, r2 = r * r, a2 = a * a;                                                         \
    const HOLE_REAL sigma = r2 + a2 * mu * mu, delta = r2 - (HOLE_REAL)2 * r + a2;                           \
    const HOLE_REAL bigA = (r2 + a2) * (r2 + a2) - a2 * delta * sine * sine;                                 \
    const HOLE_REAL rho = SQRT(sigma), rootDelta = SQRT(delta), rootA = SQRT(bigA);                          \
    const HOLE_REAL varpiAxis = rootA / rho;                                                                 \
    out->metric[0] = rho * rootDelta / rootA;                                                                \
    out->metric[1] = (HOLE_REAL)2 * a * r / bigA * varpiAxis * sine;                                         \
    out->metric[2] = varpiAxis * sine;                                                                       \
    out->metric[3] = varpiAxis;                                                                              \
    out->place[0] = rho, out->place[1] = rootDelta, out->place[2] = (HOLE_REAL)1 / r, out->place[3] = mu;    \
    out->spin[0] = a, out->spin[1] = (HOLE_REAL).995f / horizon;                                             \
    out->spin[2] = d->discHeight, out->spin[3] = d->shutter;                                                 \
    /* camera axes, scaled by the half-angles, then taken into (r, theta, phi) at the eye */               \
    HOLE_REAL fx = (HOLE_REAL)c->target.x - (HOLE_REAL)EYE(x);                                               \
    HOLE_REAL fy = (HOLE_REAL)c->target.y - (HOLE_REAL)EYE(y);                                               \
    HOLE_REAL fz = (HOLE_REAL)c->target.z - (HOLE_REAL)EYE(z);                                               \
    {                                                                                                        \
        const HOLE_REAL l = SQRT(fx * fx + fy * fy + fz * fz);                                               \
        if (l > 0) {                                                                                         \
            const HOLE_REAL inv = (HOLE_REAL)1 / l;                                                          \
            fx *= inv, fy *= inv, fz *= inv;                                                                 \
        }                                                                                                    \
    }                                                                                                        \
    const HOLE_REAL ux = c->up.x, uy = c->up.y, uz = c->up.z;                                                \
    HOLE_REAL rx = fy * uz - fz * uy, ry = fz * ux - fx * uz, rz = fx * uy - fy * ux;                        \
    {                                                                                                        \
        const HOLE_REAL l = SQRT(rx * rx + ry * ry + rz * rz);                                               \
        if (l > 0) {                                                                                         \
            const HOLE_REAL inv = (HOLE_REAL)1 / l;                                                          \
            rx *= inv, ry *= inv, rz *= inv;                                                                 \
        }                                                                                                    \
    }                                                                                                        \
    const HOLE_REAL qx = ry * fz - rz * fy, qy = rz * fx - rx * fz, qz = rx * fy - ry * fx;                  \
    const HOLE_REAL tanY = TAN((HOLE_REAL)c->fovY * (HOLE_REAL).5f);                                        \
    const HOLE_REAL tanX = tanY * (HOLE_REAL)width / (HOLE_REAL)artfMaxi(height, 1);                        \
    HOLE_REAL* const slots[4] = {out->right, out->up, out->forward, out->boost};                             \
    HOLE_REAL vel[3] = {d->velocity.x, d->velocity.y, d->velocity.z};                                        \
    HOLE_REAL speed = SQRT(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);                             \
    if (speed > (HOLE_REAL).99f) {                                                                           \
        const HOLE_REAL k = (HOLE_REAL).99f / speed;                                                         \
        vel[0] *= k, vel[1] *= k, vel[2] *= k;                                                               \
    }                                                                                                        \
    speed = speed < (HOLE_REAL).99f ? speed : (HOLE_REAL).99f;                                               \
    const HOLE_REAL vecs[4][3] = {{rx * tanX, ry * tanX, rz * tanX}, {qx * tanY, qy * tanY, qz * tanY}, {fx, fy, fz}, {vel[0], vel[1], vel[2]}}; \
    for (int k = 0; k < 4; ++k) {                                                                            \
        const HOLE_REAL* v = vecs[k];                                                                        \
        const HOLE_REAL hx = v[0] * exx + v[1] * exy + v[2] * exz;                                           \
        const HOLE_REAL hy = v[0] * ezx + v[1] * ezy + v[2] * ezz;                                           \
        const HOLE_REAL hz = v[0] * eyx + v[1] * eyy + v[2] * eyz;                                           \
        slots[k][0] = hx * er[0] + hy * er[1] + hz * er[2];                                                  \
        slots[k][1] = hx * et[0] + hy * et[1] + hz * et[2];                                                  \
        slots[k][2] = hx * ep[0] + hy * ep[1] + hz * ep[2];                                                  \
        slots[k][3] = 0;                                                                                     \
    }                                                                                                        \
    out->boost[3] = (HOLE_REAL)1 / SQRT((HOLE_REAL)1 - speed * speed);                                       \
    out->disc[0] = d->discInner, out->disc[1] = d->discOuter, out->disc[2] = d->time, out->disc[3] = d->discThickness;

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

void artfHoleFrame64Of(const ArtfBlackHoleDesc* d, const ArtfCamera* c, int width, int height, ArtfHoleFrame64* out) {
    memset(out, 0, sizeof *out);
    HOLE_FRAME_BODY(double, sqrt, tan, fabs, (double)d->spin, CAMERA_EYE)
}

/* The FP64 frame as a function of a spin and an eye position given in doubles (for derivatives). */
void artfHoleFrame64At(const ArtfBlackHoleDesc* d, const ArtfCamera* c, double spin, const double eye[3], int width, int height, ArtfHoleFrame64* out) {
    memset(out, 0, sizeof *out);
#define GIVEN_EYE(k) eye[#k[0] - 'x']
    HOLE_FRAME_BODY(double, sqrt, tan, fabs, spin, GIVEN_EYE)
#undef GIVEN_EYE
}

void artfHoleFrameLevels(const ArtfScene* s, int width, int height, int levelsCap, ArtfHoleFrame* out, uint32_t* bloomPixels) {
    const double t = artfHoleFluidTime(&s->hole);
    artfHoleFrameOf(&s->hole.desc, &s->camera, t, width, height, levelsCap, out, bloomPixels);
}

void artfHoleFrame(const ArtfScene* s, int width, int height, ArtfHoleFrame* out, uint32_t* bloomPixels) {
    artfHoleFrameLevels(s, width, height, 0, out, bloomPixels);
}

void artfHoleFrame64(const ArtfScene* s, int width, int height, ArtfHoleFrame64* out) { artfHoleFrame64Of(&s->hole.desc, &s->camera, width, height, out); }
