/* artful-hip — the tracers for a batch of views: a workgroup a tile (viewTilesKernel, at the end) or
 * persistent and work-queued (queueKernel).
 *
 * A tile is 64 rays of one view: its pixels' rays in adjacent lanes, as the frame's tile kernel has them
 * (ARTF_HIP_TILE_W x ARTF_HIP_TILE_H pixels of side^2 rays; 8x8 pixels at one ray a pixel).
 *
 * The queued kernel: a fixed grid of workgroups stays on the device until every ray of every view has
 * ended. Each wave (64 lanes, one ray a lane) takes chunks of tiles from a queue (a counter, one atomic add
 * per chunk), keeps them in a wave-private pool, and refills: once at least
 * schedule.refill of its lanes are idle, it finishes their rays (the sky's light and the write, all
 * those lanes together), then hands the idle lanes the pool's next rays, in lane order (a ballot and
 * a prefix count over the idle lanes), and sets them up.
 *
 * A wave holds the rays of one view at a time: its view is a wave-uniform value, and on each change
 * of view the wave copies the view's entry (frames, ray law, outputs: 800 bytes) into its own slot of
 * LDS, from where every lane reads it by broadcast. The per-lane state is then only the ray's own,
 * and a disc crossing or a ray's setup waits on LDS, not on a trip to L2, for the view's constants
 * (which the tile kernel has in SGPRs from its arguments). When the pool's next rays belong to
 * another view, the wave takes them once its lanes have all ended.
 *
 * There is a queue per XCD, each holding a contiguous eighth of the tiles: a wave takes from its own
 * XCD's queue first, so that the eight L2s each see one compact part of the views (and the disc tables
 * it reads), and the atomics are spread over eight counters; once its queue is empty it takes from the
 * others in turn, which balances the end.
 *
 * Per ray, the arithmetic is the tile kernel's (tracer.h: setupRay, nystromStep, the crossings,
 * skyLight, in the same order), so each view's image equals what artfHipTrace draws for its frame,
 * bit for bit in FP32 (trace_exact.hip). Not in the reference, which draws one view a frame. */
#pragma once

#include "trace_kernel.h"

namespace artf {

constexpr int kQueueGroup = 256;
constexpr uint32_t kNoRay = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

template <typename R> __device__ __forceinline__ void eyeOfView(const ArtfHipView& v, Eye<R>& e);
template <> __device__ __forceinline__ void eyeOfView<float>(const ArtfHipView& v, Eye<float>& e) { eyeFrom<float>(v.frame, e); }
template <> __device__ __forceinline__ void eyeOfView<double>(const ArtfHipView& v, Eye<double>& e) { eyeFrom64(v.frame64, e); }

/* The disc's constants as discFrom makes them from the view's frame (FP32) or frames (FP64). */
__device__ __forceinline__ void discOfView(const ArtfHipView& v, float time, DiscConst<float>& d) {
    const ArtfHoleFrame& f = v.frame;
    discOf<float>(f, f.disc[0], f.disc[1], f.disc[3], f.spin[2], time, d);
}

__device__ __forceinline__ void discOfView(const ArtfHipView& v, float time, DiscConst<double>& d) {
    const ArtfHoleFrame64& g = v.frame64;
    discOf<double>(v.frame, g.disc[0], g.disc[1], g.disc[3], g.spin[2], time, d);
}

/* The wave's view: its entry, tables and ray law (all wave-uniform). */
struct WaveView {
    const ArtfHipView* v;
    DiscTables tables;
    float uEscape;
};

/* The wave's copy of views[view] in its LDS slot, then the view's table addresses and ray law. */
__device__ __forceinline__ WaveView waveViewOf(const ArtfHipView* __restrict__ views, const ArtfHipBatchArgs& a, uint32_t view,
                                               ArtfHipView* slot, int lane) {
    static_assert(sizeof(ArtfHipView) % 8 == 0, "the view is copied in 8-byte words");
    const uint2* from = reinterpret_cast<const uint2*>(views + view);
    uint2* to = reinterpret_cast<uint2*>(slot);
    for (int i = lane; i < (int)(sizeof(ArtfHipView) / 8); i += 64) to[i] = from[i];
    __builtin_amdgcn_wave_barrier();
    WaveView w;
    w.v = slot;
    w.tables.map = a.map;
    w.tables.rings = a.rings + (size_t)view * ARTF_HIP_RING_FLOATS;
    w.tables.glow = w.tables.rings + ARTF_HOLE_MAP_HEIGHT * 2;
    w.tables.fluid = a.fluids ? a.fluids + (size_t)view * ((size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT) : a.map;
    w.uEscape = w.v->escapeRadius > 0.f ? 1.f / w.v->escapeRadius : 0.f;
    return w;
}

/* Where a ray left for the sky: the direction and drag angle at the escape. */
template <typename R> struct Away {
    R x, y, z, psi;
};

/* A lane's ray in one precision R: the state traceRay keeps across its loop. */
template <typename R> struct PlainLane {
    using Aim = R;
    RayConst<R> c;
    RayState<R> s;
    Gathered<R> acc;
    Away<R> away;

    __device__ __forceinline__ void begin(const WaveView& w, R nx, R ny, bool inRangeDrag) {
        Eye<R> e;
        eyeOfView<R>(*w.v, e);
        setupRay(e, nx, ny, c, s);
        c.inRangeDrag = inRangeDrag;
        acc.r = acc.g = acc.b = (R)0, acc.through = (R)1;
    }

    /* One step of traceRay's loop; true when the ray ends in it (sky: by escaping). */
    __device__ __forceinline__ bool step(const WaveView& w, float time, bool& sky) {
        const ArtfHipView& v = *w.v;
        RayState<R> n;
        StepLooks<R> looks;
        nystromStep(c, s, n, looks);
        if (s.n2 * n.n2 < (R)0) {
            DiscConst<R> d;
            discOfView(v, time, d);
            if (v.disc == DISC_PAINTED) crossPainted(d, c, s, n, looks, w.tables, acc);
            else crossAnalytic(d, c, s, n, looks, v.emissivity, w.tables.map, acc);
        }
        const R uEscape = (R)w.uEscape;
        const bool escaped = uEscape > n.u;
        if (escaped || n.u > uHorizonOf(v) || !(acc.through > (R).004f)) { /* where it left, only as it leaves: as traceRay */
            if (escaped) {
                const R part = quot(s.u - uEscape, s.u - n.u);
                away.x = s.n0 + part * (n.n0 - s.n0);
                away.y = s.n1 + part * (n.n1 - s.n1);
                away.z = s.n2 + part * (n.n2 - s.n2);
                away.psi = s.psi + part * (n.psi - s.psi);
                sky = true;
            }
            return true;
        }
        s = n;
        return false;
    }

    __device__ __forceinline__ R uHorizonOf(const ArtfHipView& v) const;

    __device__ __forceinline__ void finish(const WaveView& w, bool sky, float rgb[3]) {
        const float stars = w.v->frame.look[2];
        if (stars > 0.f && sky) skyLight(stars, c, away.x, away.y, away.z, away.psi, (R)w.v->frame.look[1], w.tables.map, acc);
        rgb[0] = (float)acc.r, rgb[1] = (float)acc.g, rgb[2] = (float)acc.b;
    }
};

/* The capture radius as eyeFrom / eyeFrom64 give it. */
template <> __device__ __forceinline__ float PlainLane<float>::uHorizonOf(const ArtfHipView& v) const { return v.frame.spin[1]; }
template <> __device__ __forceinline__ double PlainLane<double>::uHorizonOf(const ArtfHipView& v) const { return v.frame64.spin[1]; }

/* A lane's ray in the mixed tracer: traceRayMixed's state (FP32 steps, FP64 where they are
 * ill-conditioned). */
struct MixedLane {
    using Aim = double;
    RayConst<double> cd;
    RayConst<float> cf;
    RayState<float> s;
    Gathered<float> acc;
    Away<float> away;

    __device__ __forceinline__ void begin(const WaveView& w, double nx, double ny, bool inRangeDrag) {
        Eye<double> e;
        eyeFrom64(w.v->frame64, e);
        RayState<double> sd;
        setupRay(e, nx, ny, cd, sd);
        cf = narrowConst(cd);
        cf.inRangeDrag = inRangeDrag;
        s = castState<float>(sd);
        acc.r = acc.g = acc.b = 0.f, acc.through = 1.f;
    }

    /* traceRayMixed's step (mixedStep: FP32, with the radial half or the polar axis again in FP64 at
     * turning points and for the step that reaches the horizon) */
    __device__ __forceinline__ bool step(const WaveView& w, float time, bool& sky) {
        const ArtfHipView& v = *w.v;
        const float uHorizon = (float)v.frame64.spin[1], uEscape = w.uEscape;
        RayState<float> n;
        StepLooks<float> looks;
        mixedStep(cd, cf, s, uHorizon, n, looks);
        if (s.n2 * n.n2 < 0.f) {
            DiscConst<float> d;
            discOfView(v, time, d);
            if (v.disc == DISC_PAINTED) crossPainted(d, cf, s, n, looks, w.tables, acc);
            else crossAnalytic(d, cf, s, n, looks, v.emissivity, w.tables.map, acc);
        }
        const bool escaped = uEscape > n.u;
        if (escaped || n.u > uHorizon || !(acc.through > .004f)) { /* where it left, only as it leaves: as traceRay */
            if (escaped) {
                const float part = quot(s.u - uEscape, s.u - n.u);
                away.x = s.n0 + part * (n.n0 - s.n0);
                away.y = s.n1 + part * (n.n1 - s.n1);
                away.z = s.n2 + part * (n.n2 - s.n2);
                away.psi = s.psi + part * (n.psi - s.psi);
                sky = true;
            }
            return true;
        }
        s = n;
        return false;
    }

    __device__ __forceinline__ void finish(const WaveView& w, bool sky, float rgb[3]) {
        const float stars = w.v->frame.look[2];
        if (stars > 0.f && sky) skyLight(stars, cf, away.x, away.y, away.z, away.psi, w.v->frame.look[1], w.tables.map, acc);
        rgb[0] = acc.r, rgb[1] = acc.g, rgb[2] = acc.b;
    }
};

template <int Precision> struct LaneOf { using type = PlainLane<float>; };
template <> struct LaneOf<ARTF_PRECISION_FP64> { using type = PlainLane<double>; };
template <> struct LaneOf<ARTF_PRECISION_MIXED> { using type = MixedLane; };

/* The first tile of queue q's share of the launch's tiles. */
__device__ __forceinline__ uint32_t segmentEdge(const ArtfHipBatchArgs& a, uint32_t q) {
    return a.tileBegin + (uint32_t)(((uint64_t)(a.tileEnd - a.tileBegin) * q) / ARTF_HIP_QUEUES);
}

/* The view holding tile t: the last v with firstTile[v] <= t (a wave-uniform search). */
__device__ __forceinline__ uint32_t viewOfTile(const uint32_t* __restrict__ firstTile, int viewCount, uint32_t tile) {
    uint32_t low = 0, high = (uint32_t)viewCount;
    while (high - low > 1) {
        const uint32_t mid = (low + high) >> 1;
        if (firstTile[mid] <= tile) low = mid;
        else high = mid;
    }
    return uniform(low);
}

/* When the wave leaves a view (or ends): the last moment its rays of that view ended, into one of the
 * view's slots, the later of the two kept. One atomic per wave and view, its result not waited for. */
__device__ __forceinline__ void stampView(const ArtfHipBatchArgs& a, uint32_t view, unsigned long long& lastEnd, int lane) {
    if (!a.stamps || !lastEnd) return;
    if (lane == 0) {
        const unsigned slot = (blockIdx.x * (kQueueGroup / 64) + (threadIdx.x >> 6)) % ARTF_HIP_STAMP_SLOTS;
        atomicMax(&a.stamps[ARTF_HIP_STAMP_AT(a.viewCount, view, slot)], lastEnd);
    }
    lastEnd = 0;
}

/* views and firstTile come apart from the rest as restrict arguments: nothing the kernel writes can
 * alias them, so their wave-uniform reads are scalar loads. */
template <int Precision, bool InRangeDrag>
__global__ __launch_bounds__(kQueueGroup) __attribute__((amdgpu_waves_per_eu(wavesPerSimd(Precision)))) void queueKernel(const ArtfHipView* __restrict__ views, const uint32_t* __restrict__ firstTile,
                                                           ArtfHipBatchArgs a) {
    using Lane = typename LaneOf<Precision>::type;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull; /* the lanes before this one */
    const int refillAt = a.schedule.refill, chunk = a.schedule.chunkTiles, stride = a.schedule.stepsPerCheck;

    __shared__ ArtfHipView cached[kQueueGroup / 64];
    ArtfHipView* slot = &cached[threadIdx.x >> 6];
    Lane ray;
    bool active = false, pending = false, sky = false;
    int steps = 0;
    uint32_t outAt = 0;
    float time = 0.f;
    unsigned long long stepSum = 0, slots = 0;
    /* wave-uniform: the pool of rays taken from the queue and not yet given out, [poolNext, poolEnd);
     * the wave's view, whose rays are [viewBegin, viewEnd) */
    uint32_t poolNext = 0, poolEnd = 0, view = 0, viewBegin = 0, viewEnd = 0;
    WaveView w = waveViewOf(views, a, 0, slot, lane);
    bool drained = false;
    const uint32_t xcd = blockIdx.x % ARTF_HIP_QUEUES;
    uint32_t empty = 0; /* the queues this wave has found empty (a bit each) */
    unsigned long long lastEnd = 0; /* when rays of the wave's view last ended (wave-uniform) */

    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int idleCount = __popcll(idle);
        if (idleCount >= refillAt) {
            /* finish the rays that ended since the last refill, all together */
            const unsigned long long ended = __ballot(pending);
            if (ended) {
                if (pending) {
                    float rgb[3];
                    ray.finish(w, sky, rgb);
                    const ArtfHipView& v = *w.v;
                    v.rays[0][outAt] = rgb[0];
                    v.rays[1][outAt] = rgb[1];
                    v.rays[2][outAt] = rgb[2];
                }
                if (a.stamps) lastEnd = wall_clock64();
                pending = false;
            }
            /* the pool's next rays of the wave's view to the idle lanes, in lane order */
            const int rank = __popcll(idle & below);
            uint32_t mine = kNoRay;
            int given = 0;
            while (given < idleCount && !drained) {
                if (poolNext == poolEnd) {
                    /* a chunk from this XCD's queue, else from the next that has tiles left */
                    bool got = false;
                    for (int t = 0; t < ARTF_HIP_QUEUES && !got; ++t) {
                        const uint32_t q = (xcd + (uint32_t)t) % ARTF_HIP_QUEUES;
                        if (empty & (1u << q)) continue;
                        const uint32_t begin = segmentEdge(a, q), end = segmentEdge(a, q + 1);
                        uint32_t base = 0;
                        if (lane == 0) base = atomicAdd(&a.queue[q * ARTF_HIP_QUEUE_STRIDE], (uint32_t)chunk);
                        base = uniform(base);
                        if (base >= end - begin) {
                            empty |= 1u << q;
                            continue;
                        }
                        poolNext = (begin + base) * 64u;
                        poolEnd = min(begin + base + (uint32_t)chunk, end) * 64u;
                        got = true;
                    }
                    if (!got) {
                        drained = true;
                        break;
                    }
                }
                if (poolNext >= viewEnd || poolNext < viewBegin) {
                    /* another view's rays: taken once no lane holds a ray of this one */
                    if (given > 0 || __ballot(active)) break;
                    stampView(a, view, lastEnd, lane);
                    view = viewOfTile(firstTile, a.viewCount, poolNext >> 6);
                    viewBegin = firstTile[view] * 64u;
                    viewEnd = firstTile[view + 1] * 64u;
                    w = waveViewOf(views, a, view, slot, lane);
                }
                const int take = (int)min((uint32_t)(idleCount - given), min(poolEnd, viewEnd) - poolNext);
                if (!active && rank >= given && rank < given + take) mine = poolNext + (uint32_t)(rank - given);
                poolNext += (uint32_t)take;
                given += take;
            }
            if (mine != kNoRay) {
                const ArtfHipView& v = *w.v;
                const ArtfHoleFrame& f = v.frame;
                const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
                const PixelTile tile = pixelTileOf(side);
                const uint32_t local = (mine >> 6) - (viewBegin >> 6), ty = local / v.tilesX, tx = local - ty * v.tilesX;
                const TileRay t = tileRayOf((int)(mine & 63u), side, tile, (int)tx * tile.w, (int)ty * tile.h);
                if (t.traced && t.px < width && t.py < height) { /* (a tile's places beyond the view's edges stay idle) */
                    float dx, dy;
                    subPixel(v.jitter, t.px, t.py, t.sx, t.sy, side, dx, dy);
                    time = rayMoment(f, t.sub, side);
                    typename Lane::Aim nx, ny;
                    aimOf(f, t.px, t.py, dx, dy, nx, ny);
                    ray.begin(w, nx, ny, InRangeDrag);
                    outAt = (side == 1 ? 0u : (uint32_t)t.sub * (uint32_t)width * (uint32_t)height) + (uint32_t)t.py * (uint32_t)width + (uint32_t)t.px;
                    steps = 0, sky = false, active = true;
                }
            }
            if (!__ballot(active)) {
                if (drained) break;
                continue;
            }
        }
        const int maxSteps = w.v->maxSteps;
        for (int k = 0; k < stride; ++k) {
            if (!__ballot(active)) break;
            slots += 64;
            if (active) {
                const bool over = ray.step(w, time, sky);
                ++steps;
                if (over || steps >= maxSteps) {
                    active = false, pending = true;
                    stepSum += (unsigned long long)steps;
                }
            }
        }
    }

    stampView(a, view, lastEnd, lane);
    if (a.counters) {
        for (int offset = 32; offset > 0; offset >>= 1) stepSum += __shfl_xor(stepSum, offset);
        if (lane == 0) {
            const unsigned slot = (blockIdx.x * (kQueueGroup / 64) + (threadIdx.x >> 6)) % ARTF_HIP_COUNTER_SLOTS;
            atomicAdd(&a.counters[slot], stepSum);
            atomicAdd(&a.counters[ARTF_HIP_COUNTER_SLOTS + slot], slots);
        }
    }
}

/* The batch with no queue (schedule.workgroups = ARTF_HIP_WORKGROUP_A_TILE): a workgroup of one wave for
 * each tile, as the frame's tile kernel has, and the hardware hands the workgroups to compute units as
 * their slots free up. A tile is one view's, so the view's frame, ray law and tables are wave-uniform
 * loads (scalar, from the restrict arguments), each ray is traced by the tile kernel's own code
 * (tracePixelRay) at its registers and waves a SIMD, and a pixel's rays are summed in the wave and its
 * mean written to the view's HDR (writeTile): no pool, queue, LDS copy of the view, refill, planes of
 * rays or resolve. It traces what the queued kernel does with refill 64 (new rays only once a wave's 64
 * lanes are idle), the same tiles the same way, bit for bit; the tuner weighs the two. Workgroup b goes
 * to XCD b % 8, and each XCD takes four neighbouring tiles of every 32, as in traceKernel. */
template <int Precision, bool InRangeDrag>
__global__ __launch_bounds__(kTraceGroup) __attribute__((amdgpu_waves_per_eu(traceWavesPerSimd(Precision)))) void viewTilesKernel(
    const ArtfHipView* __restrict__ views, const uint32_t* __restrict__ firstTile, ArtfHipBatchArgs a) {
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const uint32_t tile = a.tileBegin + ((slot >> 2) * 8u + xcd) * 4u + (slot & 3u);
    if (tile >= a.tileEnd) return; /* the grid rounded up to 32: the whole wave leaves */
    const uint32_t view = viewOfTile(firstTile, a.viewCount, tile);
    const ArtfHipView& v = views[view];
    const ArtfHoleFrame& f = v.frame;
    const int lane = (int)threadIdx.x;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
    /* the view's rows of tiles top to bottom; the lane's ray in its tile, as the frame's tile kernel has it */
    const PixelTile shape = pixelTileOf(side);
    const uint32_t local = tile - firstTile[view], ty = local / v.tilesX, tx = local - ty * v.tilesX;
    const int x0 = (int)tx * shape.w, y0 = (int)ty * shape.h;
    const TileRay t = tileRayOf(lane, side, shape, x0, y0);
    const bool valid = t.traced && t.px < width && t.py < height;

    TraceLaw law;
    law.maxSteps = v.maxSteps;
    law.uEscape = v.escapeRadius > 0.f ? 1.f / v.escapeRadius : 0.f;
    law.disc = v.disc;
    law.q = v.emissivity;
    law.inRangeDrag = InRangeDrag;
    DiscTables tables;
    tables.map = a.map;
    tables.rings = a.rings + (size_t)view * ARTF_HIP_RING_FLOATS;
    tables.glow = tables.rings + ARTF_HOLE_MAP_HEIGHT * 2;
    tables.fluid = a.fluids ? a.fluids + (size_t)view * ((size_t)ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT) : a.map;

    RayResult ray = {0.f, 0.f, 0.f, 0};
    if (valid) ray = tracePixelRay<Precision>(f, v.frame64, law, tables, v.jitter, t.px, t.py, t.sx, t.sy, t.sub, side);
    /* the wave's steps, and its lane-slots as the queued kernel counts them (64 lanes a step the wave takes) */
    if (a.counters) {
        const uint32_t sum = waveSum((uint32_t)ray.steps), most = waveMax((uint32_t)ray.steps);
        if (lane == 0) {
            const unsigned at = blockIdx.x % ARTF_HIP_COUNTER_SLOTS;
            atomicAdd(&a.counters[at], (unsigned long long)sum);
            atomicAdd(&a.counters[ARTF_HIP_COUNTER_SLOTS + at], (unsigned long long)most * 64ull);
        }
    }
    if (a.stamps && lane == 0) atomicMax(&a.stamps[ARTF_HIP_STAMP_AT(a.viewCount, view, blockIdx.x % ARTF_HIP_STAMP_SLOTS)], (unsigned long long)wall_clock64());
    writeTile(ray, valid, lane, side, shape, x0, y0, t.px, t.py, width, height, v.hdr, f.image[3]);
}

/* The batch's kernel, with the drag's quotient in range where every view's spin allows it: the queued
 * one, or with ARTF_HIP_WORKGROUP_A_TILE the one with a workgroup a tile. */
template <int Precision> inline hipError_t launchQueueWith(const ArtfHipBatchArgs& a, hipStream_t stream) {
    if (a.schedule.workgroups == ARTF_HIP_WORKGROUP_A_TILE) {
        const uint32_t tiles = a.tileEnd - a.tileBegin;
        if (a.tileEnd <= a.tileBegin || tiles > (1u << 25)) return hipErrorInvalidValue;
        void (*kernel)(const ArtfHipView*, const uint32_t*, ArtfHipBatchArgs) =
            a.dragInRange ? viewTilesKernel<Precision, true> : viewTilesKernel<Precision, false>;
        hipLaunchKernelGGL(kernel, dim3((tiles + 31u) / 32u * 32u), dim3(kTraceGroup), 0, stream, a.views, a.firstTile, a);
        return hipGetLastError();
    }
    void (*kernel)(const ArtfHipView*, const uint32_t*, ArtfHipBatchArgs) =
        a.dragInRange ? queueKernel<Precision, true> : queueKernel<Precision, false>;
    hipLaunchKernelGGL(kernel, dim3((unsigned)a.schedule.workgroups), dim3(kQueueGroup), 0, stream, a.views, a.firstTile, a);
    return hipGetLastError();
}

/* Workgroups of the batch's kernel a compute unit holds (the fewer of its two forms). */
template <int Precision> inline int queueOccupancy() {
    int blocks = 0, other = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, queueKernel<Precision, false>, kQueueGroup, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&other, queueKernel<Precision, true>, kQueueGroup, 0) != hipSuccess)
        return 0;
    return blocks < other ? blocks : other;
}

} // namespace artf
