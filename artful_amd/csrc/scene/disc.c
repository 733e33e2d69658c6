/* artful-hip — the disc's painted gas and its per-ring tables, on the host.
 *
 * The tracer reads one block of floats (artful/hole.h): gas detail, per-ring speed and heat,
 * black-body colours, the sky's dust, per-ring light, the disc's dust. These must be the very bits
 * Artful 0.4.0 makes (src/scene/hole.c:10-197), since every later comparison is against pictures
 * traced through them; so the arithmetic below is the reference's, operation for operation and in the
 * same order (the library is built with -ffp-contract=off, as the reference is): lattice value noise
 * with smoothstep blending, octave sums, the strand/cloud/edge recipe of hole.c:97-119, Planck's law
 * through an exponential made of multiplications only (hole.c:55-68), and the ring tables of
 * hole.c:186-197. Each texel depends on its coordinates alone, so rows go to the thread pool in any
 * order. */
#include "scene_internal.h"

#include "../core/pool.h"

/* The lattice hash (hole.c:14-21): multiply-xor-shift of the three words. */
static uint32_t latticeMix(uint32_t x, uint32_t y, uint32_t seed) {
    uint32_t h = (x * 0x9E3779B1u) ^ (y * 0x85EBCA77u) ^ (seed * 0xC2B2AE3Du);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

static float smoothStep(float lo, float hi, float x) {
    const float t = artfClampf((x - lo) / (hi - lo), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}

/* Value noise in 0..1 that wraps every `period` cells along x (hole.c:29-39). The x lattice index
 * goes through a signed integer before the wrap, so negative cells wrap as unsigned words do. */
static float valueNoise(float x, float y, uint32_t period, uint32_t seed) {
    const float cellX = floorf(x), cellY = floorf(y);
    const float fracX = x - cellX, fracY = y - cellY;
    const uint32_t x0 = (uint32_t)(int32_t)cellX % period;
    const uint32_t x1 = (x0 + 1) % period;
    const uint32_t y0 = (uint32_t)(int32_t)cellY, y1 = y0 + 1;
    const float c00 = (float)(latticeMix(x0, y0, seed) >> 8), c10 = (float)(latticeMix(x1, y0, seed) >> 8);
    const float c01 = (float)(latticeMix(x0, y1, seed) >> 8), c11 = (float)(latticeMix(x1, y1, seed) >> 8);
    const float bx = fracX * fracX * (3.f - 2.f * fracX), by = fracY * fracY * (3.f - 2.f * fracY);
    const float near = c00 + (c10 - c00) * bx, far = c01 + (c11 - c01) * bx;
    return (near + (far - near) * by) * (1.f / 16777216.f);
}

/* Octaves of value noise, normalised to 0..1 (hole.c:42-52). */
static float octaves(float x, float y, uint32_t period, float scaleY, float falloff, float stretch, int count, uint32_t seed) {
    float sum = 0, weight = .5f, weights = 0;
    for (int o = 0; o < count; ++o) {
        sum += weight * valueNoise(x * (float)period, y * scaleY, period, seed + (uint32_t)o);
        weights += weight;
        weight *= falloff;
        period *= 2;
        scaleY *= stretch;
    }
    return sum / weights;
}

/* e^x as (1 + x/2^20)^(2^20): twenty squarings (hole.c:55-59). */
static double expBySquaring(double x) {
    double v = 1.0 + x / 1048576.0;
    for (int i = 0; i < 20; ++i) v *= v;
    return v;
}

/* Planck's radiance at 611, 549 and 464 nm (hole.c:62-68; the constant is h c / k in metre-kelvin). */
static void planckRgb(double kelvin, double* out) {
    static const double wavelength[3] = {611e-9, 549e-9, 464e-9};
    for (int c = 0; c < 3; ++c) {
        const double l = wavelength[c], l2 = l * l;
        out[c] = 1.0 / (l2 * l2 * l * (expBySquaring(.014388 / (l * kelvin)) - 1.0));
    }
}

/* A black body's colour, white-balanced at 6500 K, the largest channel 1, then the film-dye tilt of
 * green down and blue up (hole.c:72-85). */
static void blackBodyRgb(double kelvin, float* rgb) {
    double v[3], white[3], top = 0;
    planckRgb(kelvin, v);
    planckRgb(6500.0, white);
    for (int c = 0; c < 3; ++c) {
        v[c] /= white[c];
        if (v[c] > top) top = v[c];
    }
    v[1] *= .6;
    v[2] *= .72;
    for (int c = 0; c < 3; ++c) rgb[c] = (float)(v[c] / top);
}

static float ridgeOf(float v) {
    const float r = 1.f - fabsf(2.f * v - 1.f);
    return r * r;
}

/* The painted gas at (turn, rho) (hole.c:97-119): strands that wander, clouds that tear further
 * out, an outer edge that frays. */
static float paintedGas(float turn, float rho) {
    const float wander = .045f * (octaves(turn, rho, 3, 5.f, .5f, 2.f, 4, 101) - .5f) + .012f * (octaves(turn, rho, 11, 24.f, .5f, 2.f, 3, 107) - .5f);
    float along = turn + .03f * (octaves(turn, rho, 5, 9.f, .5f, 2.f, 3, 113) - .5f);
    along -= floorf(along);
    const float across = rho + wander;
    float hair = 0, weight = .5f, weights = 0, scale = 600.f;
    uint32_t period = 6;
    for (int o = 0; o < 6; ++o) {
        hair += weight * ridgeOf(valueNoise(along * (float)period, across * scale, period, 120u + (uint32_t)o));
        weights += weight;
        weight *= .62f;
        period *= 2;
        scale *= 1.75f;
    }
    hair /= weights;
    const float cloud = octaves(along, across, 7, 34.f, .55f, 2.f, 5, 131);
    const float outward = smoothStep(.85f, 1.f, rho);
    const float torn = smoothStep(.3f + .08f * outward, .7f + .08f * outward, cloud);
    const float gaps = 1.f + (torn - 1.f) * smoothStep(.1f, .6f, rho);
    const float edge = smoothStep(0.f, .02f, rho) * (1.f - smoothStep(.9f, 1.f, across));
    return edge * (.25f + 2.2f * hair) * (.04f + 1.5f * gaps);
}

static void paintRow(void* user, uint32_t y, uint32_t worker) {
    (void)worker;
    float* row = (float*)user + (size_t)y * ARTF_HOLE_MAP_WIDTH;
    const float rho = (float)y / (float)(ARTF_HOLE_MAP_HEIGHT - 1);
    const float thin = 1.f - .887f * rho; /* thinner outward, as (inner/r)^2 for the default disc */
    for (int x = 0; x < ARTF_HOLE_MAP_WIDTH; ++x) {
        float turn = (float)x / (float)ARTF_HOLE_MAP_WIDTH + rho * .09f;
        turn -= floorf(turn);
        row[x] = paintedGas(turn, rho) * thin * thin;
    }
}

/* Colours by heat, linear in kelvin from 0 (the kelvin goes through a float, as artfMaxf's argument). */
void artfDiscColors(float* colors) {
    for (int i = 0; i < ARTF_HOLE_COLORS; ++i) {
        const double kelvin = artfMaxf((float)(ARTF_HOLE_KELVIN * (double)i / (ARTF_HOLE_COLORS / 2)), 400.f);
        blackBodyRgb(kelvin, colors + i * 3);
    }
}

void artfDiscPaint(float* map) {
    ArtfPool* pool = artfPoolNew(0);
    artfPoolRun(pool, paintRow, map, ARTF_HOLE_MAP_HEIGHT);
    artfPoolFree(pool);
    /* a mean of 1 over the inner half of the rows, summed in order in doubles */
    const size_t innerTexels = (size_t)ARTF_HOLE_MAP_WIDTH * (ARTF_HOLE_MAP_HEIGHT / 2);
    double sum = 0;
    for (size_t i = 0; i < innerTexels; ++i) sum += map[i];
    const float scale = (float)((double)innerTexels / sum);
    for (size_t i = 0; i < (size_t)ARTF_HOLE_MAP_WIDTH * ARTF_HOLE_MAP_HEIGHT; ++i) map[i] *= scale;

    /* the disc's dust lanes, from a little way out */
    float* dust = map + ARTF_HOLE_DUST_AT;
    for (int y = 0; y < ARTF_HOLE_DUST_HEIGHT; ++y) {
        const float rho = (float)y / (float)(ARTF_HOLE_DUST_HEIGHT - 1);
        const float outward = smoothStep(.25f, .8f, rho);
        for (int x = 0; x < ARTF_HOLE_DUST_WIDTH; ++x) {
            const float lanes = octaves((float)x / (float)ARTF_HOLE_DUST_WIDTH, rho, 4, 90.f, .55f, 2.f, 5, 151);
            dust[y * ARTF_HOLE_DUST_WIDTH + x] = outward * smoothStep(.45f, .7f, lanes);
        }
    }

    artfDiscColors(map + ARTF_HOLE_COLORS_AT);

    /* the sky: faint dust heaped along a wandering band */
    float* sky = map + ARTF_HOLE_SKY_AT;
    for (int y = 0; y < ARTF_HOLE_SKY_HEIGHT; ++y) {
        const float height = (float)y / (float)(ARTF_HOLE_SKY_HEIGHT - 1);
        for (int x = 0; x < ARTF_HOLE_SKY_WIDTH; ++x) {
            const float turn = (float)x / (float)ARTF_HOLE_SKY_WIDTH;
            const float bent = height - .5f + .22f * (valueNoise(turn * 2.f, 3.f, 2, 41) - .5f) + .1f * (valueNoise(turn * 3.f, 5.f, 3, 42) - .5f);
            const float band = 1.f - smoothStep(0.f, .3f, fabsf(bent));
            const float cloud = octaves(turn, height, 8, 8.f, .58f, 2.f, 6, 51);
            sky[y * ARTF_HOLE_SKY_WIDTH + x] = .0006f + band * band * smoothStep(.3f, .75f, cloud) * .006f;
        }
    }
}

/* Per ring (rows even in 1/r between inner and outer): turns per unit time of a Kerr circular orbit
 * 1/(2 pi (r^1.5 + a)), the one heat of 4500 K in table units, and light 6 (r/inner)^1.5. */
void artfDiscRingTables(float* rows, float* glow, float inner, float outer, float spin) {
    for (int y = 0; y < ARTF_HOLE_MAP_HEIGHT; ++y) {
        const float rho = (float)y / (float)(ARTF_HOLE_MAP_HEIGHT - 1);
        const float r = inner / (1.f - rho * (1.f - inner / outer));
        rows[y * 2] = 1.f / (6.2831853f * (r * sqrtf(r) + spin));
        rows[y * 2 + 1] = 4500.f / (float)ARTF_HOLE_KELVIN;
        const float out = r / inner;
        glow[y] = 6.f * out * sqrtf(out);
    }
}

void artfDiscRings(float* map, float inner, float outer, float spin) { artfDiscRingTables(map + ARTF_HOLE_ROWS_AT, map + ARTF_HOLE_GLOW_AT, inner, outer, spin); }
