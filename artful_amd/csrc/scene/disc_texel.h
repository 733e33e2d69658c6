/* artful-hip — one texel of the disc map's painted tables, written once for the host (scene/disc.c,
 * threaded by rows) and the GPU (hip/disc.hip, a lane a texel): the reference's arithmetic
 * (src/scene/hole.c:10-178), operation for operation and in its order, as a function of the texel's
 * coordinates alone. */
#ifndef ARTFUL_DISC_TEXEL_H
#define ARTFUL_DISC_TEXEL_H

#include "artful/hole.h"
#include "shared_math.h"

/* Value noise in 0..1 that wraps every `period` cells along x (hole.c:29-39). The x lattice index goes
 * through a signed 32-bit integer before the wrap, as x86's truncating conversion (cvttss2si) takes the
 * reference's float-to-unsigned cast: a cell below zero (the gas near its inner edge reaches y < 0)
 * wraps as an unsigned word, where a GPU's unsigned conversion would clamp it to 0. */
ARTF_SHARED float discValueNoise(float x, float y, uint32_t period, uint32_t seed) {
    const float cellX = floorf(x), cellY = floorf(y);
    const float fracX = x - cellX, fracY = y - cellY;
    const uint32_t x0 = (uint32_t)(int32_t)cellX % period;
    const uint32_t x1 = (x0 + 1) % period;
    const uint32_t y0 = (uint32_t)(int32_t)cellY, y1 = y0 + 1;
    const float c00 = (float)(sharedLattice(x0, y0, seed) >> 8), c10 = (float)(sharedLattice(x1, y0, seed) >> 8);
    const float c01 = (float)(sharedLattice(x0, y1, seed) >> 8), c11 = (float)(sharedLattice(x1, y1, seed) >> 8);
    const float bx = fracX * fracX * (3.f - 2.f * fracX), by = fracY * fracY * (3.f - 2.f * fracY);
    const float near = c00 + (c10 - c00) * bx, far = c01 + (c11 - c01) * bx;
    return (near + (far - near) * by) * (1.f / 16777216.f);
}

/* Octaves of value noise, normalised to 0..1 (hole.c:42-52). */
ARTF_SHARED float discOctaves(float x, float y, uint32_t period, float scaleY, float falloff, float stretch, int count, uint32_t seed) {
    float sum = 0, weight = .5f, weights = 0;
    for (int o = 0; o < count; ++o) {
        sum += weight * discValueNoise(x * (float)period, y * scaleY, period, seed + (uint32_t)o);
        weights += weight;
        weight *= falloff;
        period *= 2;
        scaleY *= stretch;
    }
    return sum / weights;
}

ARTF_SHARED float discRidge(float v) {
    const float r = 1.f - fabsf(2.f * v - 1.f);
    return r * r;
}

/* The painted gas at (turn, rho) (hole.c:97-119): strands that wander, clouds that tear further out, an
 * outer edge that frays. */
ARTF_SHARED float discPaintedGas(float turn, float rho) {
    const float wander = .045f * (discOctaves(turn, rho, 3, 5.f, .5f, 2.f, 4, 101) - .5f) + .012f * (discOctaves(turn, rho, 11, 24.f, .5f, 2.f, 3, 107) - .5f);
    float along = turn + .03f * (discOctaves(turn, rho, 5, 9.f, .5f, 2.f, 3, 113) - .5f);
    along -= floorf(along);
    const float across = rho + wander;
    float hair = 0, weight = .5f, weights = 0, scale = 600.f;
    uint32_t period = 6;
    for (int o = 0; o < 6; ++o) {
        hair += weight * discRidge(discValueNoise(along * (float)period, across * scale, period, 120u + (uint32_t)o));
        weights += weight;
        weight *= .62f;
        period *= 2;
        scale *= 1.75f;
    }
    hair /= weights;
    const float cloud = discOctaves(along, across, 7, 34.f, .55f, 2.f, 5, 131);
    const float outward = sharedSmoothStep(.85f, 1.f, rho);
    const float torn = sharedSmoothStep(.3f + .08f * outward, .7f + .08f * outward, cloud);
    const float gaps = 1.f + (torn - 1.f) * sharedSmoothStep(.1f, .6f, rho);
    const float edge = sharedSmoothStep(0.f, .02f, rho) * (1.f - sharedSmoothStep(.9f, 1.f, across));
    return edge * (.25f + 2.2f * hair) * (.04f + 1.5f * gaps);
}

/* Gas texel (x, y) of the map before its normalisation: thinner outward, as (inner/r)^2 for the default
 * disc (hole.c:122-134). */
ARTF_SHARED float discGasTexel(int x, int y) {
    const float rho = (float)y / (float)(ARTF_HOLE_MAP_HEIGHT - 1);
    const float thin = 1.f - .887f * rho;
    float turn = (float)x / (float)ARTF_HOLE_MAP_WIDTH + rho * .09f;
    turn -= floorf(turn);
    return discPaintedGas(turn, rho) * thin * thin;
}

/* Dust texel (x, y) of the disc: lanes, from a little way out (hole.c:147-156). */
ARTF_SHARED float discDustTexel(int x, int y) {
    const float rho = (float)y / (float)(ARTF_HOLE_DUST_HEIGHT - 1);
    const float outward = sharedSmoothStep(.25f, .8f, rho);
    const float lanes = discOctaves((float)x / (float)ARTF_HOLE_DUST_WIDTH, rho, 4, 90.f, .55f, 2.f, 5, 151);
    return outward * sharedSmoothStep(.45f, .7f, lanes);
}

/* Sky texel (x, y): faint dust heaped along a wandering band (hole.c:166-176). */
ARTF_SHARED float discSkyTexel(int x, int y) {
    const float height = (float)y / (float)(ARTF_HOLE_SKY_HEIGHT - 1);
    const float turn = (float)x / (float)ARTF_HOLE_SKY_WIDTH;
    const float bent = height - .5f + .22f * (discValueNoise(turn * 2.f, 3.f, 2, 41) - .5f) + .1f * (discValueNoise(turn * 3.f, 5.f, 3, 42) - .5f);
    const float band = 1.f - sharedSmoothStep(0.f, .3f, fabsf(bent));
    const float cloud = discOctaves(turn, height, 8, 8.f, .58f, 2.f, 6, 51);
    return .0006f + band * band * sharedSmoothStep(.3f, .75f, cloud) * .006f;
}

#endif
