/* artful-hip — the disc's gas as a 2-D isothermal fluid on rings, on the host.
 *
 * The tracer reads one field of it (1024 cells around x 256 rings). Artful 0.4.0 runs this on the CPU
 * for every backend (src/scene/fluid.c:1-24) and the flow is chaotic, so any difference in any bit
 * grows: the scheme below is the reference's, step for step and operation for operation —
 *   - state: density S and the velocity's departure from a circular orbit (vr, w), double-buffered;
 *   - a step (fluid.c:231-241): 2 large and 14 small random eddies stirred in, in order, from one
 *     xorshift32 stream; velocities advected semi-Lagrangian and pushed by Coriolis, Kerr epicyclic
 *     swing and pressure, with the new w used at once for vr; density advected and compressed by
 *     the new divergence, drifting inward and refilling; each ring then shifted along itself by its
 *     own orbit with a monotone Catmull-Rom cubic (the FARGO idea); each ring's mean pulled toward 1;
 *   - what shows: the density times an envelope in r, torn where it is thin (fluid.c:175-185).
 * Each row of each pass reads the previous buffers only and writes its own row, so rows go to the
 * thread pool in any order and the bytes do not depend on the thread count. */
#include "scene_internal.h"

#include "../core/pool.h"

#define FW ARTF_FLUID_WIDTH
#define FH ARTF_FLUID_HEIGHT
#define FLUID_DT ARTF_FLUID_DT
#define FLUID_DPHI (6.2831853f / (float)FW) /* a cell's angle */
#define FLUID_CATCH_UP ARTF_FLUID_CATCH_UP
#define FLUID_WARM ARTF_FLUID_WARM
#define FLUID_WOUND ARTF_FLUID_WOUND

struct ArtfFluid {
    float inner, outer;
    float* density[2];
    float* radial[2];  /* vr */
    float* around[2];  /* w */
    float* field;
    int cur;
    ArtfFluidRings k;
    uint32_t rng;
    double time;
    ArtfPool* pool;
};

static uint32_t rngNext(uint32_t* rng) {
    uint32_t x = *rng;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *rng = x;
    return x;
}
static float rngUnit(uint32_t* rng) { return (float)(rngNext(rng) >> 8) * (1.f / 16777216.f); }

static float smoothStep(float lo, float hi, float x) {
    const float t = artfClampf((x - lo) / (hi - lo), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}

static uint32_t latticeMix(uint32_t x, uint32_t y, uint32_t seed) {
    uint32_t h = (x * 0x9E3779B1u) ^ (y * 0x85EBCA77u) ^ (seed * 0xC2B2AE3Du);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

/* Value noise wrapping every `period` (a power of two) along x, 0..1 (fluid.c:71-79). */
static float wrappedNoise(float x, float y, uint32_t period, uint32_t seed) {
    const float cellX = floorf(x), cellY = floorf(y), fracX = x - cellX, fracY = y - cellY;
    const uint32_t x0 = (uint32_t)cellX & (period - 1), x1 = (x0 + 1) & (period - 1), y0 = (uint32_t)cellY, y1 = y0 + 1;
    const float c00 = (float)(latticeMix(x0, y0, seed) >> 8), c10 = (float)(latticeMix(x1, y0, seed) >> 8);
    const float c01 = (float)(latticeMix(x0, y1, seed) >> 8), c11 = (float)(latticeMix(x1, y1, seed) >> 8);
    const float bx = fracX * fracX * (3.f - 2.f * fracX), by = fracY * fracY * (3.f - 2.f * fracY);
    const float near = c00 + (c10 - c00) * bx, far = c01 + (c11 - c01) * bx;
    return (near + (far - near) * by) * (1.f / 16777216.f);
}

/* Bilinear read at (x, y) in cells: wrapping around a ring, clamped across rings (fluid.c:82-92). */
static inline float readAt(const float* a, float x, float y) {
    const float yc = artfClampf(y, 0.f, (float)(FH - 1) - .001f);
    const int j0 = (int)yc;
    const float fy = yc - (float)j0;
    const float xf = floorf(x), fx = x - xf;
    const int i0 = (int)xf & (FW - 1), i1 = (i0 + 1) & (FW - 1);
    const float* lo = a + (size_t)j0 * FW;
    const float* hi = lo + FW;
    const float near = lo[i0] + (lo[i1] - lo[i0]) * fx, far = hi[i0] + (hi[i1] - hi[i0]) * fx;
    return near + (far - near) * fy;
}

/* New velocities of one ring (fluid.c:96-123). */
static void velocityPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int j = (int)row, c = f->cur;
    const float* s = f->density[c];
    const float* vr = f->radial[c];
    const float* w = f->around[c];
    float* outVr = f->radial[c ^ 1] + (size_t)j * FW;
    float* outW = f->around[c ^ 1] + (size_t)j * FW;
    const float r = f->k.r[j], invR = 1.f / r, cs2 = f->k.sound2[j], om = f->k.omega[j], sw = f->k.swing[j];
    const float cellsPerW = FLUID_DT / (r * FLUID_DPHI), rowsPerVr = FLUID_DT / f->k.rowSpan[j];
    const int jUp = artfMini(j + 1, FH - 1), jDown = artfMaxi(j - 1, 0);
    const float invSpan = 1.f / (f->k.r[jUp] - f->k.r[jDown]), invArc = 1.f / (2.f * r * FLUID_DPHI);
    const float damp = 1.f - FLUID_DT * .04f * om;
    for (int i = 0; i < FW; ++i) {
        const size_t at = (size_t)j * FW + i;
        const float fromX = (float)i - w[at] * cellsPerW, fromY = (float)j - vr[at] * rowsPerVr;
        const float v0 = readAt(vr, fromX, fromY), w0 = readAt(w, fromX, fromY);
        const float sHere = s[at], invS = 1.f / sHere;
        const float gradR = (s[(size_t)jUp * FW + i] - s[(size_t)jDown * FW + i]) * invSpan * invS;
        const float gradPhi = (s[(size_t)j * FW + ((i + 1) & (FW - 1))] - s[(size_t)j * FW + ((i - 1) & (FW - 1))]) * invArc * invS;
        const float w1 = w0 + FLUID_DT * (-sw * v0 - v0 * w0 * invR - cs2 * gradPhi);
        const float v1 = v0 + FLUID_DT * (2.f * om * w1 + w1 * w1 * invR - cs2 * gradR);
        outW[i] = artfClampf(w1 * damp, -.4f, .4f);
        outVr[i] = artfClampf(v1 * damp, -.4f, .4f);
    }
}

/* New density of one ring (fluid.c:126-149). */
static void densityPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int j = (int)row, c = f->cur;
    const float* s = f->density[c];
    const float* vr = f->radial[c ^ 1];
    const float* w = f->around[c ^ 1];
    float* out = f->density[c ^ 1] + (size_t)j * FW;
    const float r = f->k.r[j], invR = 1.f / r;
    const float cellsPerW = FLUID_DT / (r * FLUID_DPHI), rowsPerVr = FLUID_DT / f->k.rowSpan[j];
    const int jUp = artfMini(j + 1, FH - 1), jDown = artfMaxi(j - 1, 0);
    const float rUp = f->k.r[jUp], rDown = f->k.r[jDown], invSpan = 1.f / (rUp - rDown), invArc = 1.f / (2.f * r * FLUID_DPHI);
    const float inflow = -.004f, refill = FLUID_DT * .002f * f->k.omega[j];
    for (int i = 0; i < FW; ++i) {
        const size_t at = (size_t)j * FW + i;
        const float fromX = (float)i - w[at] * cellsPerW, fromY = (float)j - (vr[at] + inflow) * rowsPerVr;
        const float carried = readAt(s, fromX, fromY);
        const float divergence = (rUp * vr[(size_t)jUp * FW + i] - rDown * vr[(size_t)jDown * FW + i]) * invSpan * invR +
                                 (w[(size_t)j * FW + ((i + 1) & (FW - 1))] - w[(size_t)j * FW + ((i - 1) & (FW - 1))]) * invArc;
        const float next = carried * (1.f - FLUID_DT * artfClampf(divergence, -2.f, 2.f));
        out[i] = artfClampf(next + (1.f - next) * refill, .03f, 8.f);
    }
}

/* One ring moved along itself by `cells` with a Catmull-Rom cubic clamped to its two middle
 * neighbours, so it makes no new extremes (fluid.c:153-161). */
static void shiftRing(const float* in, float* out, float cells) {
    const float whole = floorf(cells), t = 1.f - (cells - whole), t2 = t * t, t3 = t2 * t;
    const int base = -(int)whole - 1;
    const float k0 = -.5f * t + t2 - .5f * t3, k1 = 1.f - 2.5f * t2 + 1.5f * t3, k2 = .5f * t + 2.f * t2 - 1.5f * t3, k3 = -.5f * t2 + .5f * t3;
    for (int i = 0; i < FW; ++i) {
        const float p0 = in[(i + base - 1) & (FW - 1)], p1 = in[(i + base) & (FW - 1)];
        const float p2 = in[(i + base + 1) & (FW - 1)], p3 = in[(i + base + 2) & (FW - 1)];
        out[i] = artfClampf(k0 * p0 + k1 * p1 + k2 * p2 + k3 * p3, artfMinf(p1, p2), artfMaxf(p1, p2));
    }
}

static void shiftPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int from = f->cur ^ 1;
    const size_t at = (size_t)row * FW;
    const float cells = f->k.shiftCells[row];
    shiftRing(f->density[from] + at, f->density[from ^ 1] + at, cells);
    shiftRing(f->radial[from] + at, f->radial[from ^ 1] + at, cells);
    shiftRing(f->around[from] + at, f->around[from ^ 1] + at, cells);
}

static void fieldPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const float* s = f->density[f->cur] + (size_t)row * FW;
    float* out = f->field + (size_t)row * FW;
    const float env = f->k.envelope[row], fall = f->k.falloff[row];
    for (int i = 0; i < FW; ++i) {
        const float v = s[i] * env;
        out[i] = fall * smoothStep(.1f, 1.1f, v) * artfMinf(v, 2.2f);
    }
}

/* An eddy at a random ring and cell: a whirl of velocity and a clump of gas (fluid.c:188-216). The
 * random draws are taken in the reference's order: radius, cell, size, sign, strength, clump. */
void artfFluidDrawEddy(uint32_t* rng, const ArtfFluidRings* k, float inner, float outer, bool small, ArtfFluidEddy* e) {
    const float rho = rngUnit(rng), r0 = inner + (outer - inner) * rho * rho;
    int j0 = 0;
    while (j0 < FH - 1 && k->r[j0] < r0) ++j0;
    const int i0 = (int)(rngNext(rng) & (FW - 1));
    float size;
    if (small) size = r0 * (.006f + .02f * rngUnit(rng));
    else size = r0 * (.02f + .1f * rngUnit(rng));
    const float cs = sqrtf(k->sound2[j0]);
    float strength = 0.f;
    if (!small) {
        const float sign = rngUnit(rng) < .5f ? -1.f : 1.f;
        const float amount = .3f + .7f * rngUnit(rng);
        strength = sign * cs * amount;
    }
    e->r0 = r0, e->size = size, e->strength = strength;
    e->clump = (rngUnit(rng) - .42f) * (small ? 1.5f : .5f);
    e->i0 = i0;
    e->reach = artfMini((int)(2.5f * size / (r0 * FLUID_DPHI)) + 1, FW / 2 - 1);
}

static void applyEddy(ArtfFluid* f, const ArtfFluidEddy* e) {
    float* s = f->density[f->cur];
    float* vr = f->radial[f->cur];
    float* w = f->around[f->cur];
    const float r0 = e->r0, size = e->size, strength = e->strength, clump = e->clump;
    const float invSize2 = 1.f / (size * size);
    for (int j = 0; j < FH; ++j) {
        const float dr = f->k.r[j] - r0;
        if (dr * dr > 6.25f * size * size) continue;
        for (int di = -e->reach; di <= e->reach; ++di) {
            const float dx = (float)di * f->k.r[j] * FLUID_DPHI;
            const float d2 = (dr * dr + dx * dx) * invSize2;
            if (d2 > 6.25f) continue;
            const float g = 1.f - d2 * .16f, shape = g * g * g;
            const size_t at = (size_t)j * FW + (size_t)((e->i0 + di) & (FW - 1));
            vr[at] -= strength * shape * dx / size;
            w[at] += strength * shape * dr / size;
            s[at] = artfClampf(s[at] * (1.f + clump * shape), .03f, 8.f);
        }
    }
}

static void stirEddy(ArtfFluid* f, bool small) {
    ArtfFluidEddy e;
    artfFluidDrawEddy(&f->rng, &f->k, f->inner, f->outer, small, &e);
    applyEddy(f, &e);
}

/* Each ring's mean drawn 3% of the way back to 1 (fluid.c:221-229). */
static void balancePass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    float* s = f->density[f->cur] + (size_t)row * FW;
    float sum = 0;
    for (int i = 0; i < FW; ++i) sum += s[i];
    const float mean = sum * (1.f / (float)FW), scale = 1.f + .03f * (1.f / mean - 1.f);
    for (int i = 0; i < FW; ++i) s[i] = artfClampf(s[i] * scale, .03f, 8.f);
}

static void fluidStep(ArtfFluid* f) {
    for (int i = 0; i < 2; ++i) stirEddy(f, false);
    for (int i = 0; i < 14; ++i) stirEddy(f, true);
    artfPoolRun(f->pool, velocityPass, f, FH);
    artfPoolRun(f->pool, densityPass, f, FH);
    artfPoolRun(f->pool, shiftPass, f, FH); /* back into the current buffers */
    artfPoolRun(f->pool, balancePass, f, FH);
    f->time += FLUID_DT;
}

void artfFluidFree(ArtfFluid* f) {
    if (!f) return;
    artfPoolFree(f->pool);
    for (int k = 0; k < 2; ++k) {
        artfFree(f->density[k]);
        artfFree(f->radial[k]);
        artfFree(f->around[k]);
    }
    artfFree(f->field);
    artfFree(f);
}

void artfFluidRingsOf(float inner, float outer, float spin, ArtfFluidRings* out) {
    const float k = 1.f - inner / outer;
    for (int j = 0; j < FH; ++j) {
        const float rho = (float)j / (float)(FH - 1), r = inner / (1.f - rho * k);
        out->r[j] = r;
        const float root = sqrtf(r), om = 1.f / (r * root + spin);
        out->omega[j] = om;
        /* Kerr's epicyclic frequency squared: zero at the last stable orbit */
        const float kappa2 = om * om * (1.f - 6.f / r + 8.f * spin / (r * root) - 3.f * spin * spin / (r * r));
        out->swing[j] = artfMaxf(kappa2, -.5f * om * om) / (2.f * om);
        const float cs = .03f * r * om; /* sound: the disc's thickness times the orbit's speed */
        out->sound2[j] = cs * cs;
        out->shiftCells[j] = om * FLUID_DT / FLUID_DPHI;
        out->envelope[j] = smoothStep(inner + .5f, inner + 2.2f, r) * (1.f - .86f * smoothStep(outer * .12f, outer * .8f, r));
        out->falloff[j] = sqrtf(inner / r);
    }
    for (int j = 0; j < FH; ++j)
        out->rowSpan[j] = (out->r[artfMini(j + 1, FH - 1)] - out->r[artfMaxi(j - 1, 0)]) * (j == 0 || j == FH - 1 ? 1.f : .5f);
}

ArtfFluid* artfFluidNew(float inner, float outer, float spin, double time) {
    ArtfFluid* f = ARTF_NEW(ArtfFluid);
    f->inner = inner;
    f->outer = outer;
    f->rng = ARTF_FLUID_SEED;
    f->time = time;
    f->pool = artfPoolNew(0);
    const size_t cells = (size_t)FW * FH;
    for (int k = 0; k < 2; ++k) {
        f->density[k] = artfMalloc(cells * 4);
        f->radial[k] = artfCalloc(cells, 4);
        f->around[k] = artfCalloc(cells, 4);
    }
    f->field = artfMalloc(cells * 4);

    artfFluidRingsOf(inner, outer, spin, &f->k);

    /* clumps already wound up by FLUID_WOUND of shear: each ring turned by its own orbit */
    float* s = f->density[0];
    for (int j = 0; j < FH; ++j) {
        const float rho = (float)j / (float)(FH - 1), wound = f->k.omega[j] * FLUID_WOUND / 6.2831853f;
        for (int i = 0; i < FW; ++i) {
            float turn = (float)i / (float)FW - wound;
            turn -= floorf(turn);
            float sum = 0, weight = .5f, weights = 0;
            uint32_t period = 16;
            float across = 14.f;
            for (uint32_t o = 0; o < 5; ++o) {
                sum += weight * wrappedNoise(turn * (float)period, rho * across, period, 7 + o);
                weights += weight;
                weight *= .72f;
                period *= 2;
                across *= 2.f;
            }
            const float v = sum / weights;
            s[(size_t)j * FW + i] = .06f + 3.4f * smoothStep(.34f, .68f, v) * smoothStep(.34f, .68f, v);
        }
    }
    for (int i = 0; i < 200; ++i) stirEddy(f, false);
    for (int i = 0; i < FLUID_WARM; ++i) fluidStep(f);
    f->time = time;
    artfPoolRun(f->pool, fieldPass, f, FH);
    return f;
}

bool artfFluidAdvance(ArtfFluid* f, double time) {
    if (time - f->time < FLUID_DT) {
        /* gas does not run back; far behind, the clock is simply reset */
        if (time < f->time - 8.0) f->time = time;
        return false;
    }
    int steps = (int)((time - f->time) / FLUID_DT);
    if (steps > FLUID_CATCH_UP) {
        f->time = time - (double)FLUID_CATCH_UP * FLUID_DT;
        steps = FLUID_CATCH_UP;
    }
    for (int i = 0; i < steps; ++i) fluidStep(f);
    artfPoolRun(f->pool, fieldPass, f, FH);
    return true;
}

const float* artfFluidField(const ArtfFluid* f) { return f->field; }
double artfFluidTime(const ArtfFluid* f) { return f->time; }
