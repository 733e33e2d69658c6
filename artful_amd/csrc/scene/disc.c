/* artful-hip — the disc's painted gas and its per-ring tables, on the host.
 *
 * The tracer reads one block of floats (artful/hole.h): gas detail, per-ring speed and heat,
 * black-body colours, the sky's dust, per-ring light, the disc's dust. These must be the very bits
 * Artful 0.4.0 makes (src/scene/hole.c:10-197), since every later comparison is against pictures
 * traced through them; so the arithmetic is the reference's, operation for operation and in the same
 * order (the library is built with -ffp-contract=off, as the reference is): each texel of the gas, the
 * dust and the sky is disc_texel.h's, the one source the GPU's painting (hip/disc.hip) also builds
 * from; here, Planck's law through an exponential made of multiplications only (hole.c:55-68), the
 * gas's normalisation and the ring tables of hole.c:186-197. Each texel depends on its coordinates
 * alone, so rows go to the thread pool in any order. */
#include "scene_internal.h"

#include "../core/pool.h"
#include "disc_texel.h"

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

static void paintRow(void* user, uint32_t y, uint32_t worker) {
    (void)worker;
    float* row = (float*)user + (size_t)y * ARTF_HOLE_MAP_WIDTH;
    for (int x = 0; x < ARTF_HOLE_MAP_WIDTH; ++x) row[x] = discGasTexel(x, (int)y);
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
    for (int y = 0; y < ARTF_HOLE_DUST_HEIGHT; ++y)
        for (int x = 0; x < ARTF_HOLE_DUST_WIDTH; ++x) dust[y * ARTF_HOLE_DUST_WIDTH + x] = discDustTexel(x, y);

    artfDiscColors(map + ARTF_HOLE_COLORS_AT);

    /* the sky: faint dust heaped along a wandering band */
    float* sky = map + ARTF_HOLE_SKY_AT;
    for (int y = 0; y < ARTF_HOLE_SKY_HEIGHT; ++y)
        for (int x = 0; x < ARTF_HOLE_SKY_WIDTH; ++x) sky[y * ARTF_HOLE_SKY_WIDTH + x] = discSkyTexel(x, y);
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
