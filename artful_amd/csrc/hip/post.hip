/* artful-hip — after the trace: each pixel's rays averaged, the bloom pyramid, and the tone map.
 * All three are HBM-bound streams over float planes, one texel per lane with consecutive lanes on
 * consecutive texels of a row (coalesced). Built with -ffp-contract=off: the arithmetic is the
 * reference's, operation for operation (src/software/software.c:320-386; lens.inc:364-380, 387-440),
 * so the bytes of a frame equal the reference's for the same light. */
#include "launch.h"
#include "tracer.h"

namespace artf {

__device__ __forceinline__ float minRef(float a, float b) { return b < a ? b : a; } /* artfMinf */
__device__ __forceinline__ float maxRef(float a, float b) { return a < b ? b : a; } /* artfMaxf */
__device__ __forceinline__ int clampInt(int v, int lo_, int hi_) { return v < lo_ ? lo_ : v > hi_ ? hi_ : v; }

struct Planes {
    float* p[3];
};

/* Mean of a pixel's side^2 rays, summed in the order lensTile adds them, times 1/side^2. */
__global__ __launch_bounds__(256) void resolveKernel(Planes rays, Planes hdr, int pixels, int count, float share) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    for (int c = 0; c < 3; ++c) {
        float sum = 0.f;
        for (int k = 0; k < count; ++k) sum = sum + rays.p[c][(size_t)k * pixels + p];
        hdr.p[c][p] = sum * share;
    }
}

/* Bilinear read of a level at (sx, sy) in its texels, clamped at its edges (software.c:320-328). */
__device__ __forceinline__ float levelTap(const float* level, int w, int h, float sx, float sy) {
    const float baseX = floorf(sx), baseY = floorf(sy), partX = sx - baseX, partY = sy - baseY;
    const int x0 = clampInt((int)baseX, 0, w - 1), x1 = clampInt((int)baseX + 1, 0, w - 1);
    const float* r0 = level + (size_t)clampInt((int)baseY, 0, h - 1) * w;
    const float* r1 = level + (size_t)clampInt((int)baseY + 1, 0, h - 1) * w;
    const float top = r0[x0] + (r0[x1] - r0[x0]) * partX, bottom = r1[x0] + (r1[x1] - r1[x0]) * partX;
    return top + (bottom - top) * partY;
}

struct Level {
    uint32_t at, w, h, pad;
};

/* Level 0: the mean of 2 x 2 pixels, each first held to what exceeds the threshold, at most the
 * ceiling (software.c:336-349). */
__global__ __launch_bounds__(256) void bloomFirstKernel(Planes hdr, Planes bloom, Level to, int width, int height) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= (int)to.w) return;
    const int lastX = width - 1, lastY = height - 1;
    const int y0 = min(y * 2, lastY), y1 = min(y * 2 + 1, lastY);
    const int x0 = min(x * 2, lastX), x1 = min(x * 2 + 1, lastX);
    for (int c = 0; c < 3; ++c) {
        const float* a = hdr.p[c] + (size_t)y0 * width;
        const float* b = hdr.p[c] + (size_t)y1 * width;
#define BRIGHT(v) minRef(maxRef((v) - ARTF_BLOOM_THRESHOLD, 0.f), ARTF_BLOOM_CEILING)
        bloom.p[c][to.at + (size_t)y * to.w + x] = (BRIGHT(a[x0]) + BRIGHT(a[x1]) + BRIGHT(b[x0]) + BRIGHT(b[x1])) * .25f;
#undef BRIGHT
    }
}

/* Level k from k-1: four bilinear taps a texel either way, sixteen texels (software.c:350-364). */
__global__ __launch_bounds__(256) void bloomDownKernel(Planes bloom, Level from, Level to) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= (int)to.w) return;
    const float cx = (float)x * 2.f + .5f, cy = (float)y * 2.f + .5f;
    for (int c = 0; c < 3; ++c) {
        const float* level = bloom.p[c] + from.at;
        float sum = 0;
        for (int k = 0; k < 4; ++k) sum += levelTap(level, (int)from.w, (int)from.h, cx + (k & 1 ? 1.f : -1.f), cy + (k & 2 ? 1.f : -1.f));
        bloom.p[c][to.at + (size_t)y * to.w + x] = sum * .25f;
    }
}

/* Level k += the level above, spread by four taps half a texel either way (software.c:367-386). */
__global__ __launch_bounds__(256) void bloomUpKernel(Planes bloom, Level from, Level to) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= (int)to.w) return;
    const float sy = ((float)y + .5f) * .5f - .5f, sx = ((float)x + .5f) * .5f - .5f;
    for (int c = 0; c < 3; ++c) {
        const float* level = bloom.p[c] + from.at;
        float sum = 0;
        for (int k = 0; k < 4; ++k) sum += levelTap(level, (int)from.w, (int)from.h, sx + (k & 1 ? .5f : -.5f), sy + (k & 2 ? .5f : -.5f));
        float* out = bloom.p[c] + to.at + (size_t)y * to.w + x;
        *out += sum * (.25f * ARTF_BLOOM_FADE);
    }
}

/* Narkowicz's fit to ACES, then a square root (light to what a screen shows; lens.inc:387-391). */
__device__ __forceinline__ float toneCurve(float x) {
    const float top = x * (x * 2.51f + .03f);
    const float bottom = x * (x * 2.43f + .59f) + .14f;
    const float v = top / bottom;
    return __builtin_sqrtf(v < 1.f ? v : 1.f);
}

/* HDR plus the bloom's level 0 (read at half resolution, bilinear, 90% grey), times the exposure, tone
 * mapped, dithered by a hash of the pixel and truncated to bytes: BGRA, alpha 255 (lens.inc:395-440). */
__global__ __launch_bounds__(256) void toneKernel(Planes hdr, Planes bloom, uint32_t* out, Level l0, int width, float exposure, float spread) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= width) return;
    const int bw = (int)l0.w, bh = (int)l0.h;
    const float sy = ((float)y + .5f) * .5f - .5f, baseY = floorf(sy), partY = sy - baseY;
    const int y0 = clampInt((int)baseY, 0, bh - 1), y1 = clampInt((int)baseY + 1, 0, bh - 1);
    const float sx = ((float)x + .5f) * .5f - .5f;
    const float baseX = floorf(sx), partX = sx - baseX;
    const float lastX = (float)(bw - 1);
    /* (v8fMax/v8fMin of the reference: a > b ? a : b, a < b ? a : b) */
    const float bx0 = baseX > 0.f ? baseX : 0.f, bx1 = baseX + 1.f > 0.f ? baseX + 1.f : 0.f;
    const int x0 = (int)(bx0 < lastX ? bx0 : lastX), x1 = (int)(bx1 < lastX ? bx1 : lastX);
    float glows[3];
    for (int c = 0; c < 3; ++c) {
        const float* plane = bloom.p[c] + l0.at;
        const float a = plane[y0 * bw + x0], b = plane[y0 * bw + x1];
        const float d = plane[y1 * bw + x0], e = plane[y1 * bw + x1];
        const float upper = a + (b - a) * partX, lower = d + (e - d) * partX;
        glows[c] = upper + (lower - upper) * partY;
    }
    const float grey = glows[0] * .3f + glows[1] * .59f + glows[2] * .11f;
    const size_t at = (size_t)y * width + x;
    const uint32_t noise = lensHash((uint32_t)(x + y * 8192), 7u, 13u);
    const float dither = (float)(int32_t)(noise & 255u) * (1.f / 255.f);
    uint32_t pixel = 0xFF000000u;
    for (int c = 0; c < 3; ++c) {
        const float glow = glows[c] + (grey - glows[c]) * ARTF_BLOOM_WHITE;
        const float light = hdr.p[c][at] + glow * spread;
        const float v = toneCurve(light * exposure) * 255.f + dither;
        const uint32_t byte = (uint32_t)(int32_t)(v < 255.f ? v : 255.f);
        pixel |= byte << (16 - 8 * c);
    }
    out[at] = pixel;
}

static Level levelOf(const ArtfHoleFrame& f, int k) {
    Level l = {f.bloom[k][0], f.bloom[k][1], f.bloom[k][2], 0};
    return l;
}

hipError_t launchResolve(const ArtfHoleFrame& f, float* const rays[3], float* const hdr[3], hipStream_t stream) {
    const int pixels = (int)(f.size[0] * f.size[1]), side = (int)f.size[2];
    Planes r = {{rays[0], rays[1], rays[2]}}, h = {{hdr[0], hdr[1], hdr[2]}};
    hipLaunchKernelGGL(resolveKernel, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, stream, r, h, pixels, side * side, f.image[3]);
    return hipGetLastError();
}

hipError_t launchBloom(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], hipStream_t stream) {
    const int levels = (int)f.size[3];
    if (levels < 1 || levels > ARTF_BLOOM_LEVELS) return hipErrorInvalidValue;
    Planes h = {{hdr[0], hdr[1], hdr[2]}}, b = {{bloom[0], bloom[1], bloom[2]}};
    for (int k = 0; k < levels; ++k) {
        const Level to = levelOf(f, k);
        const dim3 grid((to.w + 255) / 256, to.h);
        if (k == 0) hipLaunchKernelGGL(bloomFirstKernel, grid, dim3(256), 0, stream, h, b, to, (int)f.size[0], (int)f.size[1]);
        else hipLaunchKernelGGL(bloomDownKernel, grid, dim3(256), 0, stream, b, levelOf(f, k - 1), to);
    }
    for (int k = levels - 2; k >= 0; --k) {
        const Level to = levelOf(f, k);
        hipLaunchKernelGGL(bloomUpKernel, dim3((to.w + 255) / 256, to.h), dim3(256), 0, stream, b, levelOf(f, k + 1), to);
    }
    return hipGetLastError();
}

hipError_t launchToneMap(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], uint32_t* bgra, hipStream_t stream) {
    const int width = (int)f.size[0], height = (int)f.size[1];
    Planes h = {{hdr[0], hdr[1], hdr[2]}}, b = {{bloom[0], bloom[1], bloom[2]}};
    const float spread = f.look[3] * (1.f - ARTF_BLOOM_FADE);
    hipLaunchKernelGGL(toneKernel, dim3((unsigned)((width + 255) / 256), (unsigned)height), dim3(256), 0, stream, h, b, bgra, levelOf(f, 0), width,
                       f.image[0], spread);
    return hipGetLastError();
}

/* A batch's workspace to zero and its start stamped, by a kernel rather than a memset, so that a graph
 * that captures it clears the workspace on every replay. */
__global__ __launch_bounds__(256) void batchResetKernel(uint32_t* queue, unsigned long long* counters, unsigned long long* stamps, int stampCount) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < ARTF_HIP_QUEUES) queue[i * ARTF_HIP_QUEUE_STRIDE] = 0u;
    if (counters && i < 2 * ARTF_HIP_COUNTER_SLOTS) counters[i] = 0ull;
    if (stamps && i > 0 && i < stampCount) stamps[i] = 0ull;
    if (stamps && i == 0) stamps[0] = wall_clock64();
}

__global__ __launch_bounds__(64) void queueResetKernel(uint32_t* queue) {
    if (threadIdx.x < ARTF_HIP_QUEUES) queue[threadIdx.x * ARTF_HIP_QUEUE_STRIDE] = 0u;
}

hipError_t launchQueueReset(const ArtfHipBatchArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(queueResetKernel, dim3(1), dim3(64), 0, stream, a.queue);
    return hipGetLastError();
}

hipError_t launchBatchReset(const ArtfHipBatchArgs& a, hipStream_t stream) {
    const int stamps = 1 + a.viewCount * ARTF_HIP_STAMP_SLOTS;
    const int n = max(2 * ARTF_HIP_COUNTER_SLOTS, stamps);
    hipLaunchKernelGGL(batchResetKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a.queue, a.counters, a.stamps, stamps);
    return hipGetLastError();
}

} // namespace artf

