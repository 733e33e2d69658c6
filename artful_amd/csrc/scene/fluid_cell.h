/* artful-hip — one cell of the disc fluid's passes, written once for the host (scene/fluid.c, a ring a
 * pool job) and the GPU (hip/fluid.hip, a lane a cell): the reference's arithmetic (src/scene/fluid.c),
 * operation for operation and in its order. The flow is chaotic, so any difference in any bit would
 * grow: both sides build this one source without contraction and with IEEE division.
 *
 * A ring's constants (FluidRow) are the same for all its cells; the host works them out once a ring,
 * the GPU once a cell, by the same operations. */
#ifndef ARTFUL_FLUID_CELL_H
#define ARTFUL_FLUID_CELL_H

#include "artful/hole.h"
#include "shared_math.h"

#define FLUID_W ARTF_FLUID_WIDTH
#define FLUID_H ARTF_FLUID_HEIGHT
#define FLUID_DPHI (6.2831853f / (float)FLUID_W) /* a cell's angle */

/* Value noise wrapping every `period` (a power of two) along x, 0..1 (fluid.c:71-79); x and y are never
 * negative here. */
ARTF_SHARED float fluidWrappedNoise(float x, float y, uint32_t period, uint32_t seed) {
    const float cellX = floorf(x), cellY = floorf(y), fracX = x - cellX, fracY = y - cellY;
    const uint32_t x0 = (uint32_t)cellX & (period - 1), x1 = (x0 + 1) & (period - 1), y0 = (uint32_t)cellY, y1 = y0 + 1;
    const float c00 = (float)(sharedLattice(x0, y0, seed) >> 8), c10 = (float)(sharedLattice(x1, y0, seed) >> 8);
    const float c01 = (float)(sharedLattice(x0, y1, seed) >> 8), c11 = (float)(sharedLattice(x1, y1, seed) >> 8);
    const float bx = fracX * fracX * (3.f - 2.f * fracX), by = fracY * fracY * (3.f - 2.f * fracY);
    const float near = c00 + (c10 - c00) * bx, far = c01 + (c11 - c01) * bx;
    return (near + (far - near) * by) * (1.f / 16777216.f);
}

/* Bilinear read at (x, y) in cells: wrapping around a ring, clamped across rings (fluid.c:82-92). */
ARTF_SHARED float fluidReadAt(const float* a, float x, float y) {
    const float yc = sharedClamp(y, 0.f, (float)(FLUID_H - 1) - .001f);
    const int j0 = (int)yc;
    const float fy = yc - (float)j0;
    const float xf = floorf(x), fx = x - xf;
    const int i0 = (int)xf & (FLUID_W - 1), i1 = (i0 + 1) & (FLUID_W - 1);
    const float* lo = a + (size_t)j0 * FLUID_W;
    const float* hi = lo + FLUID_W;
    const float near = lo[i0] + (lo[i1] - lo[i0]) * fx, far = hi[i0] + (hi[i1] - hi[i0]) * fx;
    return near + (far - near) * fy;
}

/* The starting density of cell (i, j): clumps already wound up by ARTF_FLUID_WOUND of shear, each ring
 * turned by its own orbit (fluid.c:286-305). */
ARTF_SHARED float fluidSeedCell(const ArtfFluidRings* k, int i, int j) {
    const float rho = (float)j / (float)(FLUID_H - 1), wound = k->omega[j] * ARTF_FLUID_WOUND / 6.2831853f;
    float turn = (float)i / (float)FLUID_W - wound;
    turn -= floorf(turn);
    float sum = 0, weight = .5f, weights = 0;
    uint32_t period = 16;
    float across = 14.f;
    for (uint32_t o = 0; o < 5; ++o) {
        sum += weight * fluidWrappedNoise(turn * (float)period, rho * across, period, 7 + o);
        weights += weight;
        weight *= .72f;
        period *= 2;
        across *= 2.f;
    }
    const float v = sum / weights;
    return .06f + 3.4f * sharedSmoothStep(.34f, .68f, v) * sharedSmoothStep(.34f, .68f, v);
}

/* What eddy e does to the cell of ring radius rj that lies di cells along the ring from its centre and
 * dr out from it: a whirl of velocity and a clump of gas (fluid.c:188-216). The caller has found the
 * ring within the eddy's reach (dr^2 <= 6.25 size^2) and di within +-reach. */
ARTF_SHARED void fluidEddyCell(const ArtfFluidEddy* e, float dr, int di, float rj, float* s, float* vr, float* w) {
    const float size = e->size;
    const float invSize2 = 1.f / (size * size);
    const float dx = (float)di * rj * FLUID_DPHI;
    const float d2 = (dr * dr + dx * dx) * invSize2;
    if (d2 > 6.25f) return;
    const float g = 1.f - d2 * .16f, shape = g * g * g;
    *vr -= e->strength * shape * dx / size;
    *w += e->strength * shape * dr / size;
    *s = sharedClamp(*s * (1.f + e->clump * shape), .03f, 8.f);
}

/* A ring's constants for the velocity and density passes (fluid.c:96-149). */
typedef struct FluidRow {
    float r, invR, cs2, om, sw, cellsPerW, rowsPerVr, rUp, rDown, invSpan, invArc, damp, refill;
    int jUp, jDown;
} FluidRow;

ARTF_SHARED FluidRow fluidRowOf(const ArtfFluidRings* k, int j) {
    FluidRow q;
    q.r = k->r[j], q.invR = 1.f / q.r, q.cs2 = k->sound2[j], q.om = k->omega[j], q.sw = k->swing[j];
    q.cellsPerW = ARTF_FLUID_DT / (q.r * FLUID_DPHI), q.rowsPerVr = ARTF_FLUID_DT / k->rowSpan[j];
    q.jUp = j + 1 < FLUID_H - 1 ? j + 1 : FLUID_H - 1, q.jDown = j - 1 > 0 ? j - 1 : 0;
    q.rUp = k->r[q.jUp], q.rDown = k->r[q.jDown];
    q.invSpan = 1.f / (q.rUp - q.rDown), q.invArc = 1.f / (2.f * q.r * FLUID_DPHI);
    q.damp = 1.f - ARTF_FLUID_DT * .04f * q.om;
    q.refill = ARTF_FLUID_DT * .002f * q.om;
    return q;
}

/* New velocities of cell (i, j) of ring q, from the state (s, vr, w) (fluid.c:96-123): advected back
 * along the flow, then Coriolis, the Kerr epicyclic swing and pressure, the new w used at once for vr. */
ARTF_SHARED void fluidVelocityCell(const FluidRow* q, const float* s, const float* vr, const float* w, int i, int j, float* outVr, float* outW) {
    const size_t at = (size_t)j * FLUID_W + i;
    const float fromX = (float)i - w[at] * q->cellsPerW, fromY = (float)j - vr[at] * q->rowsPerVr;
    const float v0 = fluidReadAt(vr, fromX, fromY), w0 = fluidReadAt(w, fromX, fromY);
    const float sHere = s[at], invS = 1.f / sHere;
    const float gradR = (s[(size_t)q->jUp * FLUID_W + i] - s[(size_t)q->jDown * FLUID_W + i]) * q->invSpan * invS;
    const float gradPhi = (s[(size_t)j * FLUID_W + ((i + 1) & (FLUID_W - 1))] - s[(size_t)j * FLUID_W + ((i - 1) & (FLUID_W - 1))]) * q->invArc * invS;
    const float w1 = w0 + ARTF_FLUID_DT * (-q->sw * v0 - v0 * w0 * q->invR - q->cs2 * gradPhi);
    const float v1 = v0 + ARTF_FLUID_DT * (2.f * q->om * w1 + w1 * w1 * q->invR - q->cs2 * gradR);
    *outW = sharedClamp(w1 * q->damp, -.4f, .4f);
    *outVr = sharedClamp(v1 * q->damp, -.4f, .4f);
}

/* New density of cell (i, j) of ring q, from the density s and the new velocities (vr, w)
 * (fluid.c:126-149): advected with a slow inflow, compressed by the divergence, refilled toward 1. */
ARTF_SHARED float fluidDensityCell(const FluidRow* q, const float* s, const float* vr, const float* w, int i, int j) {
    const size_t at = (size_t)j * FLUID_W + i;
    const float inflow = -.004f;
    const float fromX = (float)i - w[at] * q->cellsPerW, fromY = (float)j - (vr[at] + inflow) * q->rowsPerVr;
    const float carried = fluidReadAt(s, fromX, fromY);
    const float divergence = (q->rUp * vr[(size_t)q->jUp * FLUID_W + i] - q->rDown * vr[(size_t)q->jDown * FLUID_W + i]) * q->invSpan * q->invR +
                             (w[(size_t)j * FLUID_W + ((i + 1) & (FLUID_W - 1))] - w[(size_t)j * FLUID_W + ((i - 1) & (FLUID_W - 1))]) * q->invArc;
    const float next = carried * (1.f - ARTF_FLUID_DT * sharedClamp(divergence, -2.f, 2.f));
    return sharedClamp(next + (1.f - next) * q->refill, .03f, 8.f);
}

/* Cell i of a ring moved along itself by `cells` with a Catmull-Rom cubic clamped to its two middle
 * neighbours, so it makes no new extremes (fluid.c:153-161). in: the ring's cells. */
ARTF_SHARED float fluidShiftCell(const float* in, int i, float cells) {
    const float whole = floorf(cells), t = 1.f - (cells - whole), t2 = t * t, t3 = t2 * t;
    const int base = -(int)whole - 1;
    const float k0 = -.5f * t + t2 - .5f * t3, k1 = 1.f - 2.5f * t2 + 1.5f * t3, k2 = .5f * t + 2.f * t2 - 1.5f * t3, k3 = -.5f * t2 + .5f * t3;
    const float p0 = in[(i + base - 1) & (FLUID_W - 1)], p1 = in[(i + base) & (FLUID_W - 1)];
    const float p2 = in[(i + base + 1) & (FLUID_W - 1)], p3 = in[(i + base + 2) & (FLUID_W - 1)];
    return sharedClamp(k0 * p0 + k1 * p1 + k2 * p2 + k3 * p3, sharedMin(p1, p2), sharedMax(p1, p2));
}

/* What shows of a cell of density `density` (fluid.c:175-185). */
ARTF_SHARED float fluidFieldCell(float density, float envelope, float falloff) {
    const float v = density * envelope;
    return falloff * sharedSmoothStep(.1f, 1.1f, v) * sharedMin(v, 2.2f);
}

/* A ring's balance: the scale that draws its mean 3% of the way back to 1, from its sum taken in cell
 * order (fluid.c:221-229). */
ARTF_SHARED float fluidBalanceScale(float sum) {
    const float mean = sum * (1.f / (float)FLUID_W);
    return 1.f + .03f * (1.f / mean - 1.f);
}

#endif
