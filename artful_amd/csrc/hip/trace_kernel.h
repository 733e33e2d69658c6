/* artful-hip — the tracing kernel: a workgroup of one wave traces every ray of a few neighbouring pixels,
 * a pixel's side^2 rays in side^2 adjacent lanes (an 8 x 8 tile at one ray a pixel, 4 x 4 pixels at 2 x 2
 * rays, 2 x 2 at 4 x 4, one pixel at 8 x 8), then sums each pixel's rays across its lanes in the order
 * lensTile adds them and writes the pixel's mean (lens.inc:355-383): 12 bytes a pixel, whatever the rays.
 * A wave is still 64 neighbouring rays, as many steps long as its longest, and rays end at their own
 * step: a wave's slot on its SIMD is taken by the next tile as soon as the wave ends, where the waves of
 * a larger workgroup would hold their slots until its slowest wave ends. Included by trace_exact.hip and
 * trace_fast.hip, which differ only in their contraction rule. */
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

/* The pixels of a wave: side^2 lanes a pixel, 64 / side^2 pixels (as a tileW x tileH block; with 3, 5,
 * 6 or 7 rays a side some lanes are left over and trace nothing). */
struct PixelTile {
    int w, h;
};

__host__ __device__ constexpr PixelTile pixelTileOf(int side) {
    return side == 1 ? PixelTile{8, 8} : side == 2 ? PixelTile{4, 4} : side == 3 ? PixelTile{7, 1} : side == 4 ? PixelTile{2, 2} : side == 5 ? PixelTile{2, 1} : PixelTile{1, 1};
}

template <int Precision>
__global__ __launch_bounds__(kTraceGroup) __attribute__((amdgpu_waves_per_eu(wavesPerSimd(Precision)))) void traceKernel(ArtfHipTraceArgs a) {
    stampStart(a.startStamp, a.spanStamp);
    const ArtfHoleFrame& f = a.frame;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2], count = side * side;
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const PixelTile tile = pixelTileOf(side);
    const int lane = (int)threadIdx.x;
    const int pixel = lane / count, sub = lane - pixel * count; /* the lane's pixel in the tile, and its ray */
    const int px = (int)blockIdx.x * tile.w + pixel % tile.w;
    const int py = a.rowBegin + (int)blockIdx.y * tile.h + pixel / tile.w;
    const int sx = sub % side, sy = sub / side;
    const bool valid = pixel < tile.w * tile.h && px < width && py < rowEnd;

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
    float r = ray.r, g = ray.g, b = ray.b;
    if (count > 1) {
        /* the pixel's rays in its lanes: r = r + t ray by ray from 0 (sy, then sx), then times the share,
         * in every lane of the pixel alike (count is the same for the whole wave) */
        const int first = pixel * count;
        float sr = 0.f, sg = 0.f, sb = 0.f;
        for (int k = 0; k < count; ++k) {
            sr = sr + __shfl(r, first + k);
            sg = sg + __shfl(g, first + k);
            sb = sb + __shfl(b, first + k);
        }
        r = sr * f.image[3], g = sg * f.image[3], b = sb * f.image[3];
    }
    if (!valid || sub != 0) return;
    const size_t at = (size_t)py * width + px;
    a.out[0][at] = r;
    a.out[1][at] = g;
    a.out[2][at] = b;
}

inline hipError_t launchTraceWith(void (*kernel)(ArtfHipTraceArgs), const ArtfHipTraceArgs& a, hipStream_t stream) {
    const int width = (int)a.frame.size[0], height = (int)a.frame.size[1], side = (int)a.frame.size[2];
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const int rows = rowEnd - a.rowBegin;
    if (width <= 0 || rows <= 0 || a.rowBegin < 0 || side < 1 || side > 8) return hipErrorInvalidValue;
    const PixelTile tile = pixelTileOf(side);
    const dim3 grid((unsigned)((width + tile.w - 1) / tile.w), (unsigned)((rows + tile.h - 1) / tile.h));
    if (grid.y > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, grid, dim3(kTraceGroup), 0, stream, a);
    return hipGetLastError();
}

} // namespace artf
