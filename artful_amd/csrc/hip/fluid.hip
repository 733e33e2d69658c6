/* artful-hip — the disc fluid's steps on the GPU, bit for bit the host's (scene/fluid.c, after Artful
 * 0.4.0's src/scene/fluid.c:96-241).
 *
 * The flow is chaotic, so every cell must see the very operations of the host code in the same order.
 * Built with -ffp-contract=off (no fused multiply-add), IEEE division, the host's clamp forms; each
 * pass reads the previous buffers and writes one cell per lane, as the host's rows do:
 *   - seed: the starting clumps (fluid.c:286-305): wrapped value-noise octaves, wound by the shear;
 *   - stir: a list of eddies applied in their order to each cell; a cell finds the one offset di of an
 *     eddy's sweep that lands on it, so it sees the same terms the host's loop over di gives it;
 *   - velocity, density, shift (a ring turned by its orbit with a clamped Catmull-Rom cubic): per cell;
 *   - balance: each ring's sum of 1024 floats taken by one lane in cell order (the host's order; a tree
 *     would round differently), then the ring scaled by every lane;
 *   - field: what shows.
 * Not in the reference (which steps its fluid on the CPU for every backend). */
#include <hip/hip_runtime.h>

#include "launch.h"

namespace artf {
namespace {

constexpr int FW = ARTF_FLUID_WIDTH, FH = ARTF_FLUID_HEIGHT;
#define FLUID_DPHI (6.2831853f / (float)FW)

__device__ __forceinline__ float clampOf(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }
__device__ __forceinline__ float minOf(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float maxOf(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ float smoothStep(float lo, float hi, float x) {
    const float t = clampOf((x - lo) / (hi - lo), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}

__device__ __forceinline__ uint32_t latticeMix(uint32_t x, uint32_t y, uint32_t seed) {
    uint32_t h = (x * 0x9E3779B1u) ^ (y * 0x85EBCA77u) ^ (seed * 0xC2B2AE3Du);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

__device__ float wrappedNoise(float x, float y, uint32_t period, uint32_t seed) {
    const float cellX = floorf(x), cellY = floorf(y), fracX = x - cellX, fracY = y - cellY;
    const uint32_t x0 = (uint32_t)cellX & (period - 1), x1 = (x0 + 1) & (period - 1), y0 = (uint32_t)cellY, y1 = y0 + 1;
    const float c00 = (float)(latticeMix(x0, y0, seed) >> 8), c10 = (float)(latticeMix(x1, y0, seed) >> 8);
    const float c01 = (float)(latticeMix(x0, y1, seed) >> 8), c11 = (float)(latticeMix(x1, y1, seed) >> 8);
    const float bx = fracX * fracX * (3.f - 2.f * fracX), by = fracY * fracY * (3.f - 2.f * fracY);
    const float near = c00 + (c10 - c00) * bx, far = c01 + (c11 - c01) * bx;
    return (near + (far - near) * by) * (1.f / 16777216.f);
}

/* Bilinear read at (x, y) in cells: wrapping around a ring, clamped across rings. */
__device__ __forceinline__ float readAt(const float* __restrict__ a, float x, float y) {
    const float yc = clampOf(y, 0.f, (float)(FH - 1) - .001f);
    const int j0 = (int)yc;
    const float fy = yc - (float)j0;
    const float xf = floorf(x), fx = x - xf;
    const int i0 = (int)xf & (FW - 1), i1 = (i0 + 1) & (FW - 1);
    const float* lo = a + (size_t)j0 * FW;
    const float* hi = lo + FW;
    const float near = lo[i0] + (lo[i1] - lo[i0]) * fx, far = hi[i0] + (hi[i1] - hi[i0]) * fx;
    return near + (far - near) * fy;
}

/* The cell a lane owns: lane number n is cell n % 1024 of ring n / 1024. */
__device__ __forceinline__ bool cellOf(int& i, int& j) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    i = at & (FW - 1), j = at / FW;
    return j < FH;
}

__global__ __launch_bounds__(256) void seedKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const ArtfFluidRings& k = *f.rings;
    const float rho = (float)j / (float)(FH - 1), wound = k.omega[j] * ARTF_FLUID_WOUND / 6.2831853f;
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
    const size_t at = (size_t)j * FW + i;
    f.density[at] = .06f + 3.4f * smoothStep(.34f, .68f, v) * smoothStep(.34f, .68f, v);
    f.radial[at] = 0.f, f.around[at] = 0.f;
    f.scratch[1][at] = 0.f, f.scratch[2][at] = 0.f;
}

__global__ __launch_bounds__(256) void stirKernel(ArtfHipFluid f, const ArtfFluidEddy* __restrict__ eddies, int count) {
    int i, j;
    if (!cellOf(i, j)) return;
    const float rj = f.rings->r[j];
    const size_t at = (size_t)j * FW + i;
    float s = f.density[at], vr = f.radial[at], w = f.around[at];
    for (int n = 0; n < count; ++n) {
        const ArtfFluidEddy e = eddies[n];
        const float size = e.size, dr = rj - e.r0;
        if (dr * dr > 6.25f * size * size) continue;
        const int di = ((i - e.i0 + FW / 2) & (FW - 1)) - FW / 2; /* the sweep's offset that lands here */
        if (di < -e.reach || di > e.reach) continue;
        const float invSize2 = 1.f / (size * size);
        const float dx = (float)di * rj * FLUID_DPHI;
        const float d2 = (dr * dr + dx * dx) * invSize2;
        if (d2 > 6.25f) continue;
        const float g = 1.f - d2 * .16f, shape = g * g * g;
        vr -= e.strength * shape * dx / size;
        w += e.strength * shape * dr / size;
        s = clampOf(s * (1.f + e.clump * shape), .03f, 8.f);
    }
    f.density[at] = s, f.radial[at] = vr, f.around[at] = w;
}

__global__ __launch_bounds__(256) void velocityKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const ArtfFluidRings& k = *f.rings;
    const float* __restrict__ s = f.density;
    const float* __restrict__ vr = f.radial;
    const float* __restrict__ w = f.around;
    const float r = k.r[j], invR = 1.f / r, cs2 = k.sound2[j], om = k.omega[j], sw = k.swing[j];
    const float cellsPerW = ARTF_FLUID_DT / (r * FLUID_DPHI), rowsPerVr = ARTF_FLUID_DT / k.rowSpan[j];
    const int jUp = min(j + 1, FH - 1), jDown = max(j - 1, 0);
    const float invSpan = 1.f / (k.r[jUp] - k.r[jDown]), invArc = 1.f / (2.f * r * FLUID_DPHI);
    const float damp = 1.f - ARTF_FLUID_DT * .04f * om;
    const size_t at = (size_t)j * FW + i;
    const float fromX = (float)i - w[at] * cellsPerW, fromY = (float)j - vr[at] * rowsPerVr;
    const float v0 = readAt(vr, fromX, fromY), w0 = readAt(w, fromX, fromY);
    const float sHere = s[at], invS = 1.f / sHere;
    const float gradR = (s[(size_t)jUp * FW + i] - s[(size_t)jDown * FW + i]) * invSpan * invS;
    const float gradPhi = (s[(size_t)j * FW + ((i + 1) & (FW - 1))] - s[(size_t)j * FW + ((i - 1) & (FW - 1))]) * invArc * invS;
    const float w1 = w0 + ARTF_FLUID_DT * (-sw * v0 - v0 * w0 * invR - cs2 * gradPhi);
    const float v1 = v0 + ARTF_FLUID_DT * (2.f * om * w1 + w1 * w1 * invR - cs2 * gradR);
    f.scratch[2][at] = clampOf(w1 * damp, -.4f, .4f);
    f.scratch[1][at] = clampOf(v1 * damp, -.4f, .4f);
}

__global__ __launch_bounds__(256) void densityKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const ArtfFluidRings& k = *f.rings;
    const float* __restrict__ s = f.density;
    const float* __restrict__ vr = f.scratch[1];
    const float* __restrict__ w = f.scratch[2];
    const float r = k.r[j], invR = 1.f / r;
    const float cellsPerW = ARTF_FLUID_DT / (r * FLUID_DPHI), rowsPerVr = ARTF_FLUID_DT / k.rowSpan[j];
    const int jUp = min(j + 1, FH - 1), jDown = max(j - 1, 0);
    const float rUp = k.r[jUp], rDown = k.r[jDown], invSpan = 1.f / (rUp - rDown), invArc = 1.f / (2.f * r * FLUID_DPHI);
    const float inflow = -.004f, refill = ARTF_FLUID_DT * .002f * k.omega[j];
    const size_t at = (size_t)j * FW + i;
    const float fromX = (float)i - w[at] * cellsPerW, fromY = (float)j - (vr[at] + inflow) * rowsPerVr;
    const float carried = readAt(s, fromX, fromY);
    const float divergence = (rUp * vr[(size_t)jUp * FW + i] - rDown * vr[(size_t)jDown * FW + i]) * invSpan * invR +
                             (w[(size_t)j * FW + ((i + 1) & (FW - 1))] - w[(size_t)j * FW + ((i - 1) & (FW - 1))]) * invArc;
    const float next = carried * (1.f - ARTF_FLUID_DT * clampOf(divergence, -2.f, 2.f));
    f.scratch[0][at] = clampOf(next + (1.f - next) * refill, .03f, 8.f);
}

/* Each ring turned along itself by its orbit's cells a step, from the step's second buffers back into
 * the state. */
__global__ __launch_bounds__(256) void shiftKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const float cells = f.rings->shiftCells[j];
    const float whole = floorf(cells), t = 1.f - (cells - whole), t2 = t * t, t3 = t2 * t;
    const int base = -(int)whole - 1;
    const float k0 = -.5f * t + t2 - .5f * t3, k1 = 1.f - 2.5f * t2 + 1.5f * t3, k2 = .5f * t + 2.f * t2 - 1.5f * t3, k3 = -.5f * t2 + .5f * t3;
    const size_t row = (size_t)j * FW;
    const int a0 = (i + base - 1) & (FW - 1), a1 = (i + base) & (FW - 1), a2 = (i + base + 1) & (FW - 1), a3 = (i + base + 2) & (FW - 1);
    float* const outs[3] = {f.density, f.radial, f.around};
    for (int q = 0; q < 3; ++q) {
        const float* in = f.scratch[q] + row;
        const float p0 = in[a0], p1 = in[a1], p2 = in[a2], p3 = in[a3];
        outs[q][row + i] = clampOf(k0 * p0 + k1 * p1 + k2 * p2 + k3 * p3, minOf(p1, p2), maxOf(p1, p2));
    }
}

/* A block a ring: one lane sums it in cell order, then every lane scales four cells. */
__global__ __launch_bounds__(256) void balanceKernel(ArtfHipFluid f) {
    __shared__ float scale;
    float* s = f.density + (size_t)blockIdx.x * FW;
    if (threadIdx.x == 0) {
        float sum = 0;
        for (int i = 0; i < FW; ++i) sum += s[i];
        const float mean = sum * (1.f / (float)FW);
        scale = 1.f + .03f * (1.f / mean - 1.f);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < FW; i += (int)blockDim.x) s[i] = clampOf(s[i] * scale, .03f, 8.f);
}

__global__ __launch_bounds__(256) void fieldKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const size_t at = (size_t)j * FW + i;
    const float v = f.density[at] * f.rings->envelope[j];
    f.field[at] = f.rings->falloff[j] * smoothStep(.1f, 1.1f, v) * minOf(v, 2.2f);
}

constexpr unsigned kCellBlocks = (unsigned)(FW * FH / 256);

} // namespace

hipError_t launchFluidSeed(const ArtfHipFluid& f, hipStream_t stream) {
    hipLaunchKernelGGL(seedKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    return hipGetLastError();
}

hipError_t launchFluidStir(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, int count, hipStream_t stream) {
    hipLaunchKernelGGL(stirKernel, dim3(kCellBlocks), dim3(256), 0, stream, f, eddies, count);
    return hipGetLastError();
}

hipError_t launchFluidStep(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, hipStream_t stream) {
    hipLaunchKernelGGL(stirKernel, dim3(kCellBlocks), dim3(256), 0, stream, f, eddies, ARTF_FLUID_STEP_EDDIES);
    hipLaunchKernelGGL(velocityKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    hipLaunchKernelGGL(densityKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    hipLaunchKernelGGL(shiftKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    hipLaunchKernelGGL(balanceKernel, dim3((unsigned)FH), dim3(256), 0, stream, f);
    return hipGetLastError();
}

hipError_t launchFluidField(const ArtfHipFluid& f, hipStream_t stream) {
    hipLaunchKernelGGL(fieldKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    return hipGetLastError();
}

} // namespace artf
