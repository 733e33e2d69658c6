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

/* n / d for 0 <= n < 4096 and 1 <= d <= 64, by the hardware's reciprocal: (n + 1/2) / d lies at least
 * 1/(2d) >= 1/128 from an integer, and the product is within 4096.5 x 2^-22 (1e-3) of it, so the
 * truncation is the quotient; 4 instructions where an integer division by a value the compiler does not
 * know expands to about 30. */
__device__ __forceinline__ int smallQuotient(int n, int d) {
    return (int)(((float)n + .5f) * __builtin_amdgcn_rcpf((float)d));
}

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
    const int count = side * side, turned = sub * 5; /* below 320 */
    const float moment = ((float)(turned - smallQuotient(turned, count) * count) + .5f) * f.image[3] - .5f;
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

/* The sum and the largest of a 64-lane wave's values, every lane active: an inclusive scan by DPP moves
 * within each row of 16 lanes (from 1, 2, 4 and 8 lanes below), then row 0's last lane added into row 1
 * and row 2's into row 3, then row 1's last lane into rows 2 and 3, so that lane 63 holds the whole
 * wave's; a lane whose source is outside its row, or a row the mask leaves out, takes the identity, 0.
 * About a seventh of what the same reductions cost by cross-lane reads through LDS. */
__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false); /* row_shr:1 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false); /* row_shr:2 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false); /* row_shr:4 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false); /* row_shr:8 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); /* row_bcast:15 into rows 1, 3 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); /* row_bcast:31 into rows 2, 3 */
    return (uint32_t)__builtin_amdgcn_readlane(x, 63);
}

__device__ __forceinline__ uint32_t waveMax(uint32_t v) {
    int x = (int)v; /* values below 2^31: a signed max is the unsigned one */
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane(x, 63);
}

/* The wave's step count and its lane-slots (its longest ray times its lanes with a pixel), added by
 * one lane to the counter slot of its workgroup (slots spread the atomics over many addresses). Called
 * by every lane of the wave. A ray takes at most 1,000,000 steps (the launch refuses more), so the
 * wave's 64 fit 32 bits. */
__device__ __forceinline__ void countSteps(unsigned long long* counters, int steps, bool valid) {
    if (!counters) return;
    const uint32_t sum = waveSum((uint32_t)steps), most = waveMax((uint32_t)steps);
    const uint32_t lanes = (uint32_t)__popcll(__ballot(valid));
    if ((threadIdx.x & 63) == 0 && lanes) {
        const unsigned slot = ((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x + (threadIdx.x >> 6)) % ARTF_HIP_COUNTER_SLOTS;
        atomicAdd(&counters[slot], (unsigned long long)sum);
        atomicAdd(&counters[ARTF_HIP_COUNTER_SLOTS + slot], (unsigned long long)most * (unsigned long long)lanes);
    }
}

/* The pixels of a wave: side^2 lanes a pixel, 64 / side^2 pixels (as a tileW x tileH block; with 3, 5,
 * 6 or 7 rays a side some lanes are left over and trace nothing): ARTF_HIP_TILE_W / _H. */
struct PixelTile {
    int w, h;
};

__host__ __device__ constexpr PixelTile pixelTileOf(int side) { return PixelTile{ARTF_HIP_TILE_W(side), ARTF_HIP_TILE_H(side)}; }

/* A lane's ray in its wave's tile: the pixel (px, py) from the tile's corner (x0, y0), and its sub-ray
 * (sx, sy), sub = sy * side + sx; false for the lanes left over. */
struct TileRay {
    int px, py, sx, sy, sub;
    bool traced;
};

__device__ __forceinline__ TileRay tileRayOf(int lane, int side, PixelTile tile, int x0, int y0) {
    const int count = side * side;
    const int pixel = smallQuotient(lane, count), sub = lane - pixel * count; /* the lane's pixel in the tile, and its ray */
    const int pixelY = smallQuotient(pixel, tile.w), pixelX = pixel - pixelY * tile.w;
    const int sy = smallQuotient(sub, side);
    return TileRay{x0 + pixelX, y0 + pixelY, sub - sy * side, sy, sub, pixel < tile.w * tile.h};
}

/* A wave's rays to its pixels: at one ray a pixel each lane writes its own; at side^2 the rays go through
 * LDS, then one lane a pixel and colour adds its pixel's rays in the order lensTile adds them, r = r + t
 * ray by ray from 0 (sy, then sx), and scales by the share (lens.inc:364-380): 3 x pixels lanes each with
 * a chain of `count` additions, where every lane walking all three colours' chains by cross-lane reads
 * made the wave wait on each read. Pixels at or past (width, rowEnd) are not written. Called by every
 * lane of the wave (a one-wave workgroup). */
__device__ __forceinline__ void writeTile(const RayResult& ray, bool valid, int lane, int side, PixelTile tile, int x0, int y0, int px, int py,
                                          int width, int rowEnd, float* const* out, float share) {
    const int count = side * side;
    if (count == 1) {
        if (!valid) return;
        const size_t at = (size_t)py * width + px;
        out[0][at] = ray.r;
        out[1][at] = ray.g;
        out[2][at] = ray.b;
        return;
    }
    __shared__ __align__(16) float light[3][kTraceGroup];
    light[0][lane] = ray.r;
    light[1][lane] = ray.g;
    light[2][lane] = ray.b;
    __syncthreads();
    const int pixels = tile.w * tile.h;
    if (lane >= 3 * pixels) return;
    const int colour = smallQuotient(lane, pixels), mine = lane - colour * pixels;
    const int mineY = smallQuotient(mine, tile.w);
    const int qx = x0 + mine - mineY * tile.w, qy = y0 + mineY;
    if (qx >= width || qy >= rowEnd) return;
    float* plane = colour == 0 ? out[0] : colour == 1 ? out[1] : out[2];
    const float* rays = &light[colour][mine * count];
    float sum = 0.f;
    if ((count & 3) == 0) {
        /* mine * count * 4 bytes from a 16-byte aligned array: whole float4s, four of them read before their
         * sixteen additions where there are sixteen rays left */
        const float4* quads = (const float4*)rays;
        int k = 0;
        for (; k + 4 <= count / 4; k += 4) {
            const float4 t0 = quads[k], t1 = quads[k + 1], t2 = quads[k + 2], t3 = quads[k + 3];
            sum = sum + t0.x, sum = sum + t0.y, sum = sum + t0.z, sum = sum + t0.w;
            sum = sum + t1.x, sum = sum + t1.y, sum = sum + t1.z, sum = sum + t1.w;
            sum = sum + t2.x, sum = sum + t2.y, sum = sum + t2.z, sum = sum + t2.w;
            sum = sum + t3.x, sum = sum + t3.y, sum = sum + t3.z, sum = sum + t3.w;
        }
        for (; k < count / 4; ++k) {
            const float4 t = quads[k];
            sum = sum + t.x, sum = sum + t.y, sum = sum + t.z, sum = sum + t.w;
        }
    } else {
        for (int k = 0; k < count; ++k) sum = sum + rays[k];
    }
    plane[(size_t)qy * width + qx] = sum * share;
}

/* Sub-ray (sx, sy) of pixel (px, py), sub = sy * side + sx of side^2, traced in the given precision: its
 * place in the pixel, its moment, its ray (lens.inc:364-372). */
template <int Precision>
__device__ __forceinline__ RayResult tracePixelRay(const ArtfHoleFrame& f, const ArtfHoleFrame64& f64, const TraceLaw& law, const DiscTables& tables,
                                                   uint32_t jitter, int px, int py, int sx, int sy, int sub, int side) {
    float dx, dy;
    subPixel(jitter, px, py, sx, sy, side, dx, dy);
    const float time = rayMoment(f, sub, side);
    if (Precision == ARTF_PRECISION_FP64) {
        Eye<double> e;
        eyeFrom64(f64, e);
        double nx, ny;
        aimOf(f, px, py, dx, dy, nx, ny);
        return traceRay<double>(f, e, law, tables, nx, ny, time);
    } else if (Precision == ARTF_PRECISION_MIXED) {
        double nx, ny;
        aimOf(f, px, py, dx, dy, nx, ny);
        return traceRayMixed(f, f64, law, tables, nx, ny, time);
    } else {
        Eye<float> e;
        eyeFrom<float>(f, e);
        float nx, ny;
        aimOf(f, px, py, dx, dy, nx, ny);
        return traceRay<float>(f, e, law, tables, nx, ny, time);
    }
}

/* InRangeDrag: the FP32 steps take the drag's quotient by quotInRange, for a frame whose spin allows it
 * (artfHipDragInRange); the same bits as with IEEE division, two instructions fewer for each of a step's two. */
template <int Precision, bool InRangeDrag = false>
__global__ __launch_bounds__(kTraceGroup) __attribute__((amdgpu_waves_per_eu(traceWavesPerSimd(Precision)))) void traceKernel(ArtfHipTraceArgs a) {
    stampStart(a.startStamp, a.spanStamp);
    const ArtfHoleFrame& f = a.frame;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const PixelTile tile = pixelTileOf(side);
    const int lane = (int)threadIdx.x;
    /* The hardware deals workgroup b = y * gridDim.x + x to XCD b % 8, and gridDim.x is a multiple of 32:
     * the x-th workgroup of a row is slot x / 8 of XCD x % 8, and each four slots of an XCD take four tiles
     * side by side. Their pixels then share the 64-byte pieces of each HDR row in one XCD's L2, which
     * writes each piece back once (with the round robin, four XCDs wrote a quarter of it each), and an
     * XCD's gathers into the disc's map come from fewer places: 113 MB a 1080p launch instead of 199. */
    const unsigned xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const int tileX = (int)(((slot >> 2) * 8u + xcd) * 4u + (slot & 3u)), tileY = (int)blockIdx.y;
    if (tileX * tile.w >= width) return; /* the row's tiles rounded up to 32: the whole wave leaves */
    const int x0 = tileX * tile.w, y0 = a.rowBegin + tileY * tile.h;
    const TileRay t = tileRayOf(lane, side, tile, x0, y0);
    const bool valid = t.traced && t.px < width && t.py < rowEnd;

    TraceLaw law;
    law.maxSteps = a.maxSteps > 0 ? a.maxSteps : 192;
    law.uEscape = a.escapeRadius > 0.f ? 1.f / a.escapeRadius : 0.f;
    law.disc = a.disc;
    law.q = a.emissivity;
    law.inRangeDrag = InRangeDrag;

    RayResult ray = {0.f, 0.f, 0.f, 0};
    if (valid) ray = tracePixelRay<Precision>(f, a.frame64, law, tablesOf(a.map, a.fluid), a.jitter, t.px, t.py, t.sx, t.sy, t.sub, side);
    countSteps(a.counters, ray.steps, valid);
    writeTile(ray, valid, lane, side, tile, x0, y0, t.px, t.py, width, rowEnd, a.out, f.image[3]);
}

inline hipError_t launchTraceWith(void (*kernel)(ArtfHipTraceArgs), const ArtfHipTraceArgs& a, hipStream_t stream) {
    const int width = (int)a.frame.size[0], height = (int)a.frame.size[1], side = (int)a.frame.size[2];
    const int rowEnd = a.rowEnd > 0 ? min(a.rowEnd, height) : height;
    const int rows = rowEnd - a.rowBegin;
    if (width <= 0 || rows <= 0 || a.rowBegin < 0 || side < 1 || side > 8) return hipErrorInvalidValue;
    const PixelTile tile = pixelTileOf(side);
    /* a row's tiles rounded up to a multiple of 32 workgroups (four for each of the 8 XCDs) */
    const unsigned tilesX = (unsigned)((width + tile.w - 1) / tile.w);
    const dim3 grid((tilesX + 31u) / 32u * 32u, (unsigned)((rows + tile.h - 1) / tile.h));
    if (grid.y > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, grid, dim3(kTraceGroup), 0, stream, a);
    return hipGetLastError();
}

} // namespace artf
