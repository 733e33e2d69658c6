/* artful-hip — after the trace: the bloom pyramid and the tone map (and, for a batch, each pixel's rays
 * averaged, which the tile tracer does in its own lanes).
 * All three are HBM-bound streams over float planes, one texel per lane with consecutive lanes on
 * consecutive texels of a row (coalesced). Built with -ffp-contract=off: the arithmetic is the
 * reference's, operation for operation (src/software/software.c:320-386; lens.inc:364-380, 387-440),
 * so the bytes of a frame equal the reference's for the same light. */
#include "launch.h"
#include "trace_kernel.h"
#include "tracer.h"

namespace artf {

__device__ __forceinline__ float minRef(float a, float b) { return b < a ? b : a; } /* artfMinf */
__device__ __forceinline__ float maxRef(float a, float b) { return a < b ? b : a; } /* artfMaxf */
__device__ __forceinline__ int clampInt(int v, int lo_, int hi_) { return v < lo_ ? lo_ : v > hi_ ? hi_ : v; }

struct Planes {
    float* p[3];
};

/* Every kernel here is launched with workgroups of one wave. In a frame they run beside the next
 * frame's trace, whose one-wave workgroups take each wave slot as it frees: a workgroup of several
 * waves would wait for that many slots free at once on one compute unit, in practice for the trace to
 * drain, and the next frame's trace would then wait for it. */
constexpr int kPost = 64;

/* The post kernels of a frame run beside the next frame's trace, whose waves (96 VGPRs, 5 a SIMD) leave 32
 * VGPRs a SIMD: a wave that needs more waits for the trace to stop dispatching, and the frame's chain of
 * post launches behind it. So these take at most 32 (gfx950 reads this attribute as half the registers of
 * its unified file: 16 here is 32). */
#define ARTF_POST_REGISTERS __attribute__((amdgpu_num_vgpr(16)))

/* Mean of a pixel's side^2 rays, summed in the order lensTile adds them, times 1/side^2. */
__global__ __launch_bounds__(kPost) void resolveKernel(Planes rays, Planes hdr, int pixels, int count, float share) {
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

/* The trace's step counters (2 x ARTF_HIP_COUNTER_SLOTS words, written before this kernel on its
 * stream) summed into settled[0..1] in host memory, with system-scope stores the host sees once the
 * stream is synchronised, and cleared for the next frame's trace; zeros where the trace did not count.
 * The first workgroup (one wave) of the kernel that follows the trace calls it, before anything can
 * return. */
struct Settle {
    unsigned long long* counters; /* NULL: not counted */
    unsigned long long* settled;  /* NULL: nothing to do */
};

__device__ void settleCounters(const Settle& z) {
    static_assert(kPost == 64, "the counters are summed by one wave");
    if (!z.settled || blockIdx.x != 0 || blockIdx.y != 0) return;
    unsigned long long* counters = z.counters;
    const int t = (int)threadIdx.x;
    for (int half = 0; half < 2; ++half) {
        unsigned long long sum = 0;
        for (int i = t; counters && i < ARTF_HIP_COUNTER_SLOTS; i += kPost) {
            sum += counters[half * ARTF_HIP_COUNTER_SLOTS + i];
            counters[half * ARTF_HIP_COUNTER_SLOTS + i] = 0ull;
        }
        for (int lanes = 32; lanes > 0; lanes >>= 1) sum += __shfl_xor(sum, lanes);
        if (t == 0) __hip_atomic_store(z.settled + half, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

/* Level 0: the mean of 2 x 2 pixels, each first held to what exceeds the threshold, at most the
 * ceiling (software.c:336-349). */
__global__ __launch_bounds__(kPost) ARTF_POST_REGISTERS void bloomFirstKernel(Planes hdr, Planes bloom, Level to, int width, int height, unsigned long long* stamp,
                                                        Settle z) {
    stampStart(stamp);
    settleCounters(z);
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

/* Level k's texel (x, y) from level k-1: four bilinear taps a texel either way, sixteen texels
 * (software.c:350-364). level: the plane of level k-1, in memory or in LDS. */
__device__ __forceinline__ float downTexel(const float* level, const Level& from, int x, int y) {
    const float cx = (float)x * 2.f + .5f, cy = (float)y * 2.f + .5f;
    float sum = 0;
    for (int k = 0; k < 4; ++k) sum += levelTap(level, (int)from.w, (int)from.h, cx + (k & 1 ? 1.f : -1.f), cy + (k & 2 ? 1.f : -1.f));
    return sum * .25f;
}

/* What level k's texel (x, y) gains from level k+1 on the way up: four taps half a texel either way
 * (software.c:367-386). level: the plane of level k+1. */
__device__ __forceinline__ float upGain(const float* level, const Level& from, int x, int y) {
    const float sy = ((float)y + .5f) * .5f - .5f, sx = ((float)x + .5f) * .5f - .5f;
    float sum = 0;
    for (int k = 0; k < 4; ++k) sum += levelTap(level, (int)from.w, (int)from.h, sx + (k & 1 ? .5f : -.5f), sy + (k & 2 ? .5f : -.5f));
    return sum * (.25f * ARTF_BLOOM_FADE);
}

__global__ __launch_bounds__(kPost) ARTF_POST_REGISTERS void bloomDownKernel(Planes bloom, Level from, Level to, unsigned long long* stamp) {
    stampStart(stamp);
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= (int)to.w) return;
    for (int c = 0; c < 3; ++c) bloom.p[c][to.at + (size_t)y * to.w + x] = downTexel(bloom.p[c] + from.at, from, x, y);
}

__global__ __launch_bounds__(kPost) ARTF_POST_REGISTERS void bloomUpKernel(Planes bloom, Level from, Level to, unsigned long long* stamp) {
    stampStart(stamp);
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= (int)to.w) return;
    for (int c = 0; c < 3; ++c) {
        float* out = bloom.p[c] + to.at + (size_t)y * to.w + x;
        *out += upGain(bloom.p[c] + from.at, from, x, y);
    }
}

/* The small end of the pyramid in one workgroup: level `first` (as the launches before left it) is taken
 * into LDS, levels first+1..last are built down and levels last-1..first+1 brought up there, with the
 * arithmetic of bloomDownKernel and bloomUpKernel, and stored: the pyramid the launches level by level
 * leave, in one launch instead of 2 (last - first) - 1. Levels first..last are packed one after
 * another; together they hold at most kSmallFloats / 3 texels, and first+1..last at most kSmallWork.
 * It runs only where nothing runs beside it (a frame on one stream, artfHipBloom): beside a trace the
 * pyramid goes level by level in one-wave workgroups, since this one's 128 KB of LDS and four waves would
 * wait for the trace to drain. Four waves share each level's work (with one wave it took 27 us a frame
 * on average over the GPU tests' frames, 45 us at 1080p). */
constexpr int kSmallThreads = 256;
constexpr int kSmallFloats = 32768; /* 128 KB of the compute unit's 160 KB */
constexpr int kSmallWork = 1024;

struct Pyramid {
    Level level[ARTF_BLOOM_LEVELS];
};

__global__ __launch_bounds__(kSmallThreads) void bloomSmallKernel(Planes bloom, Pyramid p, int first, int last, unsigned long long* stamp) {
    stampStart(stamp);
    __shared__ float lds[kSmallFloats];
    const int t = (int)threadIdx.x;
    const uint32_t base = p.level[first].at;
    const uint32_t n = p.level[last].at + p.level[last].w * p.level[last].h - base; /* texels a plane */
    const Level& top = p.level[first];
    for (int c = 0; c < 3; ++c)
        for (uint32_t i = t; i < top.w * top.h; i += kSmallThreads) lds[c * n + i] = bloom.p[c][top.at + i];
    __syncthreads();
    for (int k = first + 1; k <= last; ++k) {
        const Level from = p.level[k - 1], to = p.level[k];
        for (uint32_t i = t; i < to.w * to.h; i += kSmallThreads) {
            const int x = (int)(i % to.w), y = (int)(i / to.w);
            for (int c = 0; c < 3; ++c) lds[c * n + (to.at - base) + i] = downTexel(lds + c * n + (from.at - base), from, x, y);
        }
        __syncthreads();
    }
    for (int k = last - 1; k > first; --k) {
        const Level from = p.level[k + 1], to = p.level[k];
        for (uint32_t i = t; i < to.w * to.h; i += kSmallThreads) {
            const int x = (int)(i % to.w), y = (int)(i / to.w);
            for (int c = 0; c < 3; ++c) {
                float* out = lds + c * n + (to.at - base) + i;
                *out += upGain(lds + c * n + (from.at - base), from, x, y);
            }
        }
        __syncthreads();
    }
    const uint32_t skip = top.w * top.h; /* level `first` is as it came */
    for (int c = 0; c < 3; ++c)
        for (uint32_t i = skip + t; i < n; i += kSmallThreads) bloom.p[c][base + i] = lds[c * n + i];
}

/* Narkowicz's fit to ACES, then a square root (light to what a screen shows; lens.inc:387-391). */
__device__ __forceinline__ float toneCurve(float x) {
    const float top = x * (x * 2.51f + .03f);
    const float bottom = x * (x * 2.43f + .59f) + .14f;
    const float v = top / bottom;
    return __builtin_sqrtf(v < 1.f ? v : 1.f);
}

/* Where the tone map stamps the GPU's clock (each NULL for none). */
struct ToneStamps {
    unsigned long long* start;     /* as the kernel starts */
    unsigned long long* alsoStart; /* the same instant, a second stage's boundary */
    unsigned long long* end;       /* as the last workgroup (in launch order) ends */
};

/* HDR plus the bloom's level 0 (read at half resolution, bilinear, 90% grey), times the exposure, tone
 * mapped, dithered by a hash of the pixel and truncated to bytes: BGRA, alpha 255 (lens.inc:395-440). */
__device__ __forceinline__ void tonePixel(const Planes& hdr, const Planes& bloom, uint32_t* out, const Level& l0, int width, float exposure,
                                          float spread, int x, int y) {
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

/* kToneRun pixels of a row a lane, kPost apart (coalesced): fewer, longer workgroups. */
constexpr int kToneRun = 4;

__global__ __launch_bounds__(kPost) ARTF_POST_REGISTERS void toneKernel(Planes hdr, Planes bloom, uint32_t* out, Level l0, int width, float exposure, float spread,
                                                  ToneStamps stamps) {
    stampStart(stamps.start, stamps.alsoStart);
    const int y = (int)blockIdx.y;
    for (int k = 0, x = (int)blockIdx.x * kPost * kToneRun + (int)threadIdx.x; k < kToneRun && x < width; ++k, x += kPost)
        tonePixel(hdr, bloom, out, l0, width, exposure, spread, x, y);
    /* the stage's end: when the last workgroup in launch order is done (its lane 0 always has a pixel) */
    if (stamps.end && blockIdx.x == gridDim.x - 1 && blockIdx.y == gridDim.y - 1 && threadIdx.x == 0)
        __hip_atomic_store(stamps.end, (unsigned long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

static Level levelOf(const ArtfHoleFrame& f, int k) {
    Level l = {f.bloom[k][0], f.bloom[k][1], f.bloom[k][2], 0};
    return l;
}

hipError_t launchResolve(const ArtfHoleFrame& f, float* const rays[3], float* const hdr[3], hipStream_t stream) {
    const int pixels = (int)(f.size[0] * f.size[1]), side = (int)f.size[2];
    Planes r = {{rays[0], rays[1], rays[2]}}, h = {{hdr[0], hdr[1], hdr[2]}};
    hipLaunchKernelGGL(resolveKernel, dim3((unsigned)((pixels + kPost - 1) / kPost)), dim3(kPost), 0, stream, r, h, pixels, side * side, f.image[3]);
    return hipGetLastError();
}

static dim3 rowsOf(const Level& l) { return dim3((l.w + kPost - 1) / kPost, l.h); }

static uint32_t texelsFrom(const ArtfHoleFrame& f, int k) {
    const int last = (int)f.size[3] - 1;
    return f.bloom[last][0] + f.bloom[last][1] * f.bloom[last][2] - f.bloom[k][0];
}

/* The level bloomSmallKernel starts from: the first whose levels to the last fit its LDS and whose
 * levels below it fit its work; the last level when none does (then it is not launched). */
static int smallFirst(const ArtfHoleFrame& f) {
    const int last = (int)f.size[3] - 1;
    for (int k = 0; k < last; ++k)
        if ((uint64_t)texelsFrom(f, k) * 3u <= (uint64_t)kSmallFloats && texelsFrom(f, k + 1) <= (uint32_t)kSmallWork) return k;
    return last;
}

/* The pyramid above level 0, which is in place (software.c:419-426): levels 1..s down one launch each,
 * the small end in one workgroup, and levels s..0 back up one launch each (s the small end's first
 * level; without smallEnd, every level one launch each way). stamp: given to the first launch; returns
 * whether there was one. */
static bool bloomRest(const ArtfHoleFrame& f, const Planes& b, hipStream_t stream, unsigned long long* stamp, bool smallEnd) {
    const int last = (int)f.size[3] - 1, small = smallEnd ? smallFirst(f) : last;
    bool any = false;
    auto first = [&]() {
        unsigned long long* at = any ? nullptr : stamp;
        any = true;
        return at;
    };
    for (int k = 1; k <= small; ++k) {
        const Level to = levelOf(f, k);
        hipLaunchKernelGGL(bloomDownKernel, rowsOf(to), dim3(kPost), 0, stream, b, levelOf(f, k - 1), to, first());
    }
    if (small < last) {
        Pyramid p = {};
        for (int k = 0; k <= last; ++k) p.level[k] = levelOf(f, k);
        hipLaunchKernelGGL(bloomSmallKernel, dim3(1), dim3(kSmallThreads), 0, stream, b, p, small, last, first());
    }
    for (int k = small < last ? small : last - 1; k >= 0; --k) {
        const Level to = levelOf(f, k);
        hipLaunchKernelGGL(bloomUpKernel, rowsOf(to), dim3(kPost), 0, stream, b, levelOf(f, k + 1), to, first());
    }
    return any;
}

hipError_t launchBloom(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], hipStream_t stream) {
    const int levels = (int)f.size[3];
    if (levels < 1 || levels > ARTF_BLOOM_LEVELS) return hipErrorInvalidValue;
    Planes h = {{hdr[0], hdr[1], hdr[2]}}, b = {{bloom[0], bloom[1], bloom[2]}};
    const Level l0 = levelOf(f, 0);
    hipLaunchKernelGGL(bloomFirstKernel, rowsOf(l0), dim3(kPost), 0, stream, h, b, l0, (int)f.size[0], (int)f.size[1], (unsigned long long*)nullptr,
                       Settle{});
    bloomRest(f, b, stream, nullptr, true);
    return hipGetLastError();
}

static float spreadOf(const ArtfHoleFrame& f) { return f.look[3] * (1.f - ARTF_BLOOM_FADE); }

static dim3 pixelRows(const ArtfHoleFrame& f) { return dim3((f.size[0] + kPost * kToneRun - 1) / (kPost * kToneRun), f.size[1]); }

hipError_t launchToneMap(const ArtfHoleFrame& f, float* const hdr[3], float* const bloom[3], uint32_t* bgra, hipStream_t stream) {
    Planes h = {{hdr[0], hdr[1], hdr[2]}}, b = {{bloom[0], bloom[1], bloom[2]}};
    hipLaunchKernelGGL(toneKernel, pixelRows(f), dim3(kPost), 0, stream, h, b, bgra, levelOf(f, 0), (int)f.size[0], f.image[0], spreadOf(f), ToneStamps{});
    return hipGetLastError();
}

/* A frame of more than ARTF_HIP_RAYS_APART rays is traced in bands of whole rows of tiles, a launch each, one
 * after another on the trace's stream, each of at least ARTF_HIP_RAYS_A_BAND rays: as the reference's Vulkan
 * backend sends such a frame out (vulkan.c:1729-1755), so that no launch holds the GPU for long (here ~6 ms
 * a band, against 46 ms for a 3840x2160 frame of 64 rays a pixel in one launch; each band's last waves
 * drain alone, 1% of such a frame in 9 bands, 2% in 17). Rays do not depend on each other, so the image is
 * the same bits; the first band stamps the trace's start. */
static hipError_t launchTraceBands(ArtfHipTraceArgs t, hipStream_t stream) {
    const int width = (int)t.frame.size[0], height = (int)t.frame.size[1], side = (int)t.frame.size[2];
    const unsigned long long perRow = (unsigned long long)width * (unsigned long long)(side * side);
    if (perRow * (unsigned long long)height <= ARTF_HIP_RAYS_APART || side < 1 || side > 8)
        return t.precision == ARTF_PRECISION_FP32 ? launchTraceExact(t, stream) : launchTraceFast(t, stream);
    const int tileRows = pixelTileOf(side).h;
    const unsigned long long perTileRow = perRow * (unsigned long long)tileRows;
    const int band = (int)((ARTF_HIP_RAYS_A_BAND + perTileRow - 1) / perTileRow) * tileRows;
    for (int y = 0; y < height; y += band) {
        ArtfHipTraceArgs b = t;
        b.rowBegin = y, b.rowEnd = min(y + band, height);
        if (y > 0) b.startStamp = b.spanStamp = nullptr;
        const hipError_t e = b.precision == ARTF_PRECISION_FP32 ? launchTraceExact(b, stream) : launchTraceFast(b, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launchFrame(const ArtfHipFrameArgs& a, hipStream_t stream) {
    const ArtfHoleFrame& f = a.trace.frame;
    unsigned long long* r = a.report;
    ArtfHipTraceArgs t = a.trace;
    t.startStamp = r ? r + 2 : nullptr;
    t.rowBegin = 0, t.rowEnd = 0; /* a frame is traced whole */
    const hipError_t traced = launchTraceBands(t, stream);
    if (traced != hipSuccess) return traced;
    Planes h = {{a.hdr[0], a.hdr[1], a.hdr[2]}}, b = {{a.bloom[0], a.bloom[1], a.bloom[2]}};
    const Level l0 = levelOf(f, 0);
    const Settle z = {t.counters, r};
    /* the rest on the second stream, once the trace is done: the first can take the next frame's trace */
    hipStream_t post = stream;
    hipError_t e;
    if (a.postStream) {
        post = (hipStream_t)a.postStream;
        if ((e = hipEventRecord((hipEvent_t)a.handoff, stream)) != hipSuccess || (e = hipStreamWaitEvent(post, (hipEvent_t)a.handoff, 0)) != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(bloomFirstKernel, rowsOf(l0), dim3(kPost), 0, post, h, b, l0, (int)f.size[0], (int)f.size[1], r ? r + 3 : nullptr, z);
    /* beside a trace (a second stream) the small end goes level by level: its one-wave kernel needs more
     * registers than a SIMD running the trace has left; alone, one launch instead of 2 (levels - s) - 1 */
    const bool rest = bloomRest(f, b, post, r ? r + 4 : nullptr, post == stream);
    const ToneStamps ts = r ? ToneStamps{r + 5, rest ? nullptr : r + 4, r + 6} : ToneStamps{};
    hipLaunchKernelGGL(toneKernel, pixelRows(f), dim3(kPost), 0, post, h, b, a.bgra, l0, (int)f.size[0], f.image[0], spreadOf(f), ts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.postStream && a.done) return hipEventRecord((hipEvent_t)a.done, post);
    return hipSuccess;
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
    const int stamps = (int)ARTF_HIP_STAMP_WORDS(a.viewCount);
    const int n = max(2 * ARTF_HIP_COUNTER_SLOTS, stamps);
    hipLaunchKernelGGL(batchResetKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a.queue, a.counters, a.stamps, stamps);
    return hipGetLastError();
}

} // namespace artf

