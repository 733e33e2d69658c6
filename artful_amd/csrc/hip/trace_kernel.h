/* artful-hip — the tracing kernel: a workgroup of one wave traces an 8 x 8 tile of pixels (neighbouring
 * rays, which take similar paths, share the wave) for one of the pixel's side^2 rays (grid z). One wave
 * a workgroup because rays end at their own step: a wave's slot on its SIMD is taken by the next tile as
 * soon as the wave ends, where the waves of a larger workgroup hold their slots until its slowest wave
 * ends. Included by trace_exact.hip and trace_fast.hip, which differ only in their contraction rule. */
#pragma once

#include "launch.h"
#include "mixed.h"
#include "tracer.h"

namespace artf {

/* A ray's place in its pixel: the reference's stratified grid (lens.inc:367), or with a seed, a
 * jittered point in the same stratum. */
__device__ __forceinline__ void subPixel(uint32_t seed, int px, int py, int sx, int sy, int side, float& dx, float& dy) {
    if (!seed) {
        dx = ((float)sx + .5f) / (float)side;
        dy = ((float)sy + .5f) / (float)side;
        return;
    }
    const uint32_t h = lensHash((uint32_t)px ^ (seed * 0x9E3779B1u), (uint32_t)py, (uint32_t)(sy * side + sx) + seed);
    const float jx = ((float)(h & 0xFFFFu) + .5f) * (1.f / 65536.f), jy = ((float)(h >> 16) + .5f) * (1.f / 65536.f);
    dx = ((float)sx + jx) / (float)side;
    dy = ((float)sy + jy) / (float)side;
}

/* The moment of sub-ray sub of side^2 within the shutter (lens.inc:371). */
__device__ __forceinline__ float rayMoment(const ArtfHoleFrame& f, int sub, int side) {
    const float moment = ((float)((sub * 5) % (side * side)) + .5f) * f.image[3] - .5f;
    return f.disc[2] + f.spin[3] * moment;
}

/* A ray's place in -1..1 across the image: in FP32 as lensTile forms it (the pixel's group of 8,
 * plus the offset, then the lane), in FP64 directly. */
__device__ __forceinline__ void aimOf(const ArtfHoleFrame& f, int px, int py, float dx, float dy, float& nx, float& ny) {
    nx = ((float)(px & 7) + ((float)(px & ~7) + dx)) * f.image[1] - 1.f;
    ny = 1.f - ((float)py + dy) * f.image[2];
}

__device__ __forceinline__ void aimOf(const ArtfHoleFrame& f, int px, int py, float dx, float dy, double& nx, double& ny) {
    nx = ((double)px + (double)dx) * (2.0 / (double)f.size[0]) - 1.0;
    ny = 1.0 - ((double)py + (double)dy) * (2.0 / (double)f.size[1]);
}

/* The wave's step count and its lane-slots (its longest ray times its lanes with a pixel), added by
 * one lane to the counter slot of its workgroup (slots spread the atomics over many addresses). */
__device__ __forceinline__ void countSteps(unsigned long long* counters, int steps, bool valid) {
    if (!counters) return;
    unsigned long long sum = (unsigned long long)steps;
    int most = steps, lanes = valid ? 1 : 0;
    for (int offset = 32; offset > 0; offset >>= 1) {
        sum += __shfl_xor(sum, offset);
        most = max(most, __shfl_xor(most, offset));
        lanes += __shfl_xor(lanes, offset);
    }
    if ((threadIdx.x & 63) == 0 && lanes) {
        const unsigned slot = ((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x + (threadIdx.x >> 6)) % ARTF_HIP_COUNTER_SLOTS;
        atomicAdd(&counters[slot], sum);
        atomicAdd(&counters[ARTF_HIP_COUNTER_SLOTS + slot], (unsigned long long)most * (unsigned long long)lanes);
    }
}

template <int Precision>
__global__ __launch_bounds__(kTraceGroup) __attribute__((amdgpu_waves_per_eu(wavesPerSimd(Precision)))) void traceKernel(ArtfHipTraceArgs a) {
    stampStart(a.startStamp, a.spanStamp);
    const ArtfHoleFrame& f = a.frame;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const int lane = (int)threadIdx.x;
    const int px = (int)blockIdx.x * kTraceTile + (lane & 7);
    const int py = a.rowBegin + (int)blockIdx.y * kTraceTile + (lane >> 3);
    const int sub = blockIdx.z, sx = sub % side, sy = sub / side;
    const bool valid = px < width && py < rowEnd;

    TraceLaw law;
    law.maxSteps = a.maxSteps > 0 ? a.maxSteps : 192;
    law.uEscape = a.escapeRadius > 0.f ? 1.f / a.escapeRadius : 0.f;
    law.disc = a.disc;
    law.q = a.emissivity;

    RayResult ray = {0.f, 0.f, 0.f, 0};
    if (valid) {
        float dx, dy;
        subPixel(a.jitter, px, py, sx, sy, side, dx, dy);
        const float time = rayMoment(f, sub, side);
        if (Precision == ARTF_PRECISION_FP64) {
            Eye<double> e;
            eyeFrom64(a.frame64, e);
            double nx, ny;
            aimOf(f, px, py, dx, dy, nx, ny);
            ray = traceRay<double>(f, e, law, a.map, a.fluid, nx, ny, time);
        } else if (Precision == ARTF_PRECISION_MIXED) {
            double nx, ny;
            aimOf(f, px, py, dx, dy, nx, ny);
            ray = traceRayMixed(f, a.frame64, law, a.map, a.fluid, nx, ny, time);
        } else {
            Eye<float> e;
            eyeFrom<float>(f, e);
            float nx, ny;
            aimOf(f, px, py, dx, dy, nx, ny);
            ray = traceRay<float>(f, e, law, a.map, a.fluid, nx, ny, time);
        }
    }
    countSteps(a.counters, ray.steps, valid);
    if (!valid) return;
    const size_t at = (side == 1 ? 0 : (size_t)sub * width * height) + (size_t)py * width + px;
    a.out[0][at] = ray.r;
    a.out[1][at] = ray.g;
    a.out[2][at] = ray.b;
}

inline hipError_t launchTraceWith(void (*kernel)(ArtfHipTraceArgs), const ArtfHipTraceArgs& a, hipStream_t stream) {
    const int width = (int)a.frame.size[0], height = (int)a.frame.size[1], side = (int)a.frame.size[2];
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const int rows = rowEnd - a.rowBegin;
    if (width <= 0 || rows <= 0 || a.rowBegin < 0 || side < 1 || side > 8) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((width + kTraceTile - 1) / kTraceTile), (unsigned)((rows + kTraceTile - 1) / kTraceTile), (unsigned)(side * side));
    hipLaunchKernelGGL(kernel, grid, dim3(kTraceGroup), 0, stream, a);
    return hipGetLastError();
}

} // namespace artf
