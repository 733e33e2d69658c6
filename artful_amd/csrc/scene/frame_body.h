/* artful-hip — the eye's frame and the Kerr metric there, as one body for float and double frames
 * (Artful 0.4.0 src/scene/hole.c:273-356). Included by hole.c (the reference's float frame, its
 * arithmetic order kept so that the 368 bytes come out bit for bit) and by frame64.c (the same
 * geometry in doubles). Expects in scope: d (the clamped ArtfBlackHoleDesc), c (the ArtfCamera), width,
 * height and out (a frame with the fields of ArtfHoleFrame). */
#ifndef ARTFUL_FRAME_BODY_H
#define ARTFUL_FRAME_BODY_H

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

#endif
