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
 * Each cell's update is fluid_cell.h's, the one source the GPU's passes (hip/fluid.hip) also build
 * from; here a ring is a pool job. Each row of each pass reads the previous buffers only and writes its
 * own row, so rows go to the thread pool in any order and the bytes do not depend on the thread count. */
#include "scene_internal.h"

#include "../core/pool.h"
#include "fluid_cell.h"

#define FW FLUID_W
#define FH FLUID_H
#define FLUID_DT ARTF_FLUID_DT
#define FLUID_CATCH_UP ARTF_FLUID_CATCH_UP
#define FLUID_WARM ARTF_FLUID_WARM

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

/* New velocities of one ring (fluid.c:96-123). */
static void velocityPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int j = (int)row, c = f->cur;
    const FluidRow q = fluidRowOf(&f->k, j);
    float* outVr = f->radial[c ^ 1] + (size_t)j * FW;
    float* outW = f->around[c ^ 1] + (size_t)j * FW;
    for (int i = 0; i < FW; ++i) fluidVelocityCell(&q, f->density[c], f->radial[c], f->around[c], i, j, outVr + i, outW + i);
}

/* New density of one ring (fluid.c:126-149). */
static void densityPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int j = (int)row, c = f->cur;
    const FluidRow q = fluidRowOf(&f->k, j);
    float* out = f->density[c ^ 1] + (size_t)j * FW;
    for (int i = 0; i < FW; ++i) out[i] = fluidDensityCell(&q, f->density[c], f->radial[c ^ 1], f->around[c ^ 1], i, j);
}

/* Each ring moved along itself by its orbit, back into the current buffers (fluid.c:153-161). */
static void shiftPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const int from = f->cur ^ 1;
    const size_t at = (size_t)row * FW;
    const float cells = f->k.shiftCells[row];
    float* const ins[3] = {f->density[from] + at, f->radial[from] + at, f->around[from] + at};
    float* const outs[3] = {f->density[from ^ 1] + at, f->radial[from ^ 1] + at, f->around[from ^ 1] + at};
    for (int q = 0; q < 3; ++q)
        for (int i = 0; i < FW; ++i) outs[q][i] = fluidShiftCell(ins[q], i, cells);
}

static void fieldPass(void* user, uint32_t row, uint32_t worker) {
    (void)worker;
    ArtfFluid* f = user;
    const float* s = f->density[f->cur] + (size_t)row * FW;
    float* out = f->field + (size_t)row * FW;
    for (int i = 0; i < FW; ++i) out[i] = fluidFieldCell(s[i], f->k.envelope[row], f->k.falloff[row]);
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
    for (int j = 0; j < FH; ++j) {
        const float dr = f->k.r[j] - e->r0;
        if (dr * dr > 6.25f * e->size * e->size) continue;
        for (int di = -e->reach; di <= e->reach; ++di) {
            const size_t at = (size_t)j * FW + (size_t)((e->i0 + di) & (FW - 1));
            fluidEddyCell(e, dr, di, f->k.r[j], s + at, vr + at, w + at);
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
    const float scale = fluidBalanceScale(sum);
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
        out->envelope[j] = sharedSmoothStep(inner + .5f, inner + 2.2f, r) * (1.f - .86f * sharedSmoothStep(outer * .12f, outer * .8f, r));
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

    /* clumps already wound up by ARTF_FLUID_WOUND of shear: each ring turned by its own orbit */
    float* s = f->density[0];
    for (int j = 0; j < FH; ++j)
        for (int i = 0; i < FW; ++i) s[(size_t)j * FW + i] = fluidSeedCell(&f->k, i, j);
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
