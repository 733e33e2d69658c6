/* artful-hip — the disc map's painted tables on the GPU, bit for bit the host's (scene/disc.c, after
 * Artful 0.4.0's src/scene/hole.c:10-178): the gas (2048 x 1024), normalised to a mean of 1 over its
 * inner half; the disc's dust lanes (1024 x 256); the sky's dust (1024 x 512). The colour table and the
 * ring tables are small and stay on the host.
 *
 * Built with -ffp-contract=off and IEEE division: each texel is the host's operations in the host's
 * order, from its coordinates alone. Two places need care to give the host's bits:
 *   - a lattice cell below zero (the gas near its inner edge reaches y < 0): the host's float-to-
 *     unsigned conversion is x86's 32-bit signed truncation (cvttss2si), so it is written here as
 *     (uint32_t)(int32_t), not as the GPU's unsigned convert, which would clamp to 0;
 *   - the normalising sum over 1,048,576 texels is taken in doubles in texel order, as the host's
 *     loop does: the whole block stages 4096 texels at a time in LDS and one lane adds them in order
 *     (about 3 ms; a tree would round differently).
 * Not in the reference (which paints on the CPU, 0.6 s). */
#include <hip/hip_runtime.h>

#include "launch.h"

namespace artf {
namespace {

constexpr int MW = ARTF_HOLE_MAP_WIDTH, MH = ARTF_HOLE_MAP_HEIGHT;
constexpr int DW = ARTF_HOLE_DUST_WIDTH, DH = ARTF_HOLE_DUST_HEIGHT;
constexpr int SW = ARTF_HOLE_SKY_WIDTH, SH = ARTF_HOLE_SKY_HEIGHT;

__device__ __forceinline__ float clampOf(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }

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

/* Value noise in 0..1 wrapping every `period` cells along x (hole.c:29-39). */
__device__ float valueNoise(float x, float y, uint32_t period, uint32_t seed) {
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

__device__ float octaves(float x, float y, uint32_t period, float scaleY, float falloff, float stretch, int count, uint32_t seed) {
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

__device__ __forceinline__ float ridgeOf(float v) {
    const float r = 1.f - fabsf(2.f * v - 1.f);
    return r * r;
}

/* The painted gas at (turn, rho) (hole.c:97-119). */
__device__ float paintedGas(float turn, float rho) {
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

__global__ __launch_bounds__(256) void gasKernel(float* map) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= MW * MH) return;
    const int x = at % MW, y = at / MW;
    const float rho = (float)y / (float)(MH - 1);
    const float thin = 1.f - .887f * rho;
    float turn = (float)x / (float)MW + rho * .09f;
    turn -= floorf(turn);
    map[at] = paintedGas(turn, rho) * thin * thin;
}

/* The sum of the inner half's texels in doubles, in texel order (disc.c artfDiscPaint), and the gas
 * scaled to make their mean 1. One block: its lanes stage a run of texels in LDS, lane 0 adds them in
 * order; then every lane scales (one block moving 8 MB: ~0.1 ms beside the sum's ~3 ms). */
constexpr int kRun = 4096;

__global__ __launch_bounds__(256) void gasScaleKernel(float* map) {
    __shared__ float run[kRun];
    __shared__ float scale;
    const int inner = MW * (MH / 2);
    double sum = 0;
    for (int base = 0; base < inner; base += kRun) {
        for (int i = (int)threadIdx.x; i < kRun; i += (int)blockDim.x) run[i] = map[base + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < kRun; ++i) sum += run[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) scale = (float)((double)inner / sum);
    __syncthreads();
    const float k = scale;
    for (int i = (int)threadIdx.x; i < MW * MH; i += (int)blockDim.x) map[i] *= k;
}

__global__ __launch_bounds__(256) void dustKernel(float* dust) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= DW * DH) return;
    const int x = at % DW, y = at / DW;
    const float rho = (float)y / (float)(DH - 1);
    const float outward = smoothStep(.25f, .8f, rho);
    const float lanes = octaves((float)x / (float)DW, rho, 4, 90.f, .55f, 2.f, 5, 151);
    dust[at] = outward * smoothStep(.45f, .7f, lanes);
}

__global__ __launch_bounds__(256) void skyKernel(float* sky) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= SW * SH) return;
    const int x = at % SW, y = at / SW;
    const float height = (float)y / (float)(SH - 1);
    const float turn = (float)x / (float)SW;
    const float bent = height - .5f + .22f * (valueNoise(turn * 2.f, 3.f, 2, 41) - .5f) + .1f * (valueNoise(turn * 3.f, 5.f, 3, 42) - .5f);
    const float band = 1.f - smoothStep(0.f, .3f, fabsf(bent));
    const float cloud = octaves(turn, height, 8, 8.f, .58f, 2.f, 6, 51);
    sky[at] = .0006f + band * band * smoothStep(.3f, .75f, cloud) * .006f;
}

unsigned blocksFor(int n) { return (unsigned)((n + 255) / 256); }

} // namespace

hipError_t launchDiscPaint(float* map, hipStream_t stream) {
    hipLaunchKernelGGL(gasKernel, dim3(blocksFor(MW * MH)), dim3(256), 0, stream, map);
    hipLaunchKernelGGL(gasScaleKernel, dim3(1), dim3(256), 0, stream, map);
    hipLaunchKernelGGL(dustKernel, dim3(blocksFor(DW * DH)), dim3(256), 0, stream, map + ARTF_HOLE_DUST_AT);
    hipLaunchKernelGGL(skyKernel, dim3(blocksFor(SW * SH)), dim3(256), 0, stream, map + ARTF_HOLE_SKY_AT);
    return hipGetLastError();
}

} // namespace artf
