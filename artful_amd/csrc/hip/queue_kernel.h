/* artful-hip — the persistent, work-queued tracer for a batch of views.
 *
 * A fixed grid of workgroups stays on the device until every ray of every view has ended. Each wave
 * (64 lanes, one ray a lane) takes tiles of 8 x 8 rays from one queue (a single counter, one atomic
 * add per chunk of tiles), keeps them in a wave-private pool, and refills: once at least
 * schedule.refill of its lanes are idle, it finishes their rays (the sky's light and the write, all
 * those lanes together), then hands the idle lanes the pool's next rays, in lane order (a ballot and
 * a prefix count over the idle lanes), and sets them up. Long rays near the photon sphere then hold
 * one lane, not a whole wave: the stepping loop stays full while rays of 50 and of 5000 steps mix.
 *
 * Per ray, the arithmetic is the tile kernel's (tracer.h: setupRay, nystromStep, the crossings,
 * skyLight, in the same order), so each view's image equals what artfHipTrace draws for its frame,
 * bit for bit in FP32 (trace_exact.hip). What a view differs in (its frame, disc tables, ray law,
 * output) is read from its ArtfHipView in global memory, per lane, where the tile kernel has it in
 * scalar registers: lanes of one wave may hold rays of different views. Not in the reference, which
 * draws one view a frame. */
#pragma once

#include "trace_kernel.h"

namespace artf {

constexpr int kQueueGroup = 256;
constexpr uint32_t kNoRay = 0xFFFFFFFFu;

/* The view, sub-ray and pixel of ray k of the batch: tile k / 64 of the queue, place k % 64 in its
 * 8 x 8 block. False for a place beyond the view's right or bottom edge. */
struct RayPlace {
    int view, sub, px, py;
};

__device__ __forceinline__ bool placeOf(const ArtfHipBatchArgs& a, uint32_t ray, RayPlace& p) {
    const uint32_t tile = ray >> 6;
    int low = 0, high = a.viewCount; /* firstTile[low] <= tile < firstTile[high] */
    while (high - low > 1) {
        const int mid = (low + high) >> 1;
        if (a.firstTile[mid] <= tile) low = mid;
        else high = mid;
    }
    const ArtfHipView& v = a.views[low];
    const uint32_t local = tile - a.firstTile[low], perSub = v.tilesX * v.tilesY;
    const uint32_t sub = local / perSub, rest = local - sub * perSub;
    const uint32_t ty = rest / v.tilesX, tx = rest - ty * v.tilesX;
    p.view = low, p.sub = (int)sub;
    p.px = (int)(tx * 8u + (ray & 7u)), p.py = (int)(ty * 8u + ((ray >> 3) & 7u));
    return p.px < (int)v.frame.size[0] && p.py < (int)v.frame.size[1];
}

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

__device__ __forceinline__ DiscTables tablesOfView(const ArtfHipView& v, const float* map) {
    DiscTables t;
    t.map = map, t.rings = v.rings, t.glow = v.glow, t.fluid = v.fluid;
    return t;
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
    R uHorizon, uEscape;
    Away<R> away;

    __device__ __forceinline__ void begin(const ArtfHipView& v, R nx, R ny) {
        Eye<R> e;
        eyeOfView<R>(v, e);
        setupRay(e, nx, ny, c, s);
        acc.r = acc.g = acc.b = (R)0, acc.through = (R)1;
        uHorizon = e.uHorizon;
        uEscape = (R)(v.escapeRadius > 0.f ? 1.f / v.escapeRadius : 0.f);
    }

    /* One step of traceRay's loop; true when the ray ends in it (sky: by escaping). */
    __device__ __forceinline__ bool step(const ArtfHipView* views, int view, const float* map, float time, bool& sky) {
        RayState<R> n;
        StepLooks<R> looks;
        nystromStep(c, s, n, looks);
        if (s.n2 * n.n2 < (R)0) {
            const ArtfHipView& v = views[view];
            DiscConst<R> d;
            discOfView(v, time, d);
            if (v.disc == DISC_PAINTED) crossPainted(d, c, s, n, looks, tablesOfView(v, map), acc);
            else crossAnalytic(d, c, s, n, looks, v.emissivity, map, acc);
        }
        const bool escaped = uEscape > n.u;
        if (escaped) {
            const R part = quot(s.u - uEscape, s.u - n.u);
            away.x = s.n0 + part * (n.n0 - s.n0);
            away.y = s.n1 + part * (n.n1 - s.n1);
            away.z = s.n2 + part * (n.n2 - s.n2);
            away.psi = s.psi + part * (n.psi - s.psi);
            sky = true;
        }
        if (escaped || n.u > uHorizon || !(acc.through > (R).004f)) return true;
        s = n;
        return false;
    }

    __device__ __forceinline__ void finish(const ArtfHipView& v, const float* map, bool sky, float rgb[3]) {
        const float stars = v.frame.look[2];
        if (stars > 0.f && sky) skyLight(stars, c, away.x, away.y, away.z, away.psi, (R)v.frame.look[1], map, acc);
        rgb[0] = (float)acc.r, rgb[1] = (float)acc.g, rgb[2] = (float)acc.b;
    }
};

/* A lane's ray in the mixed tracer: traceRayMixed's state (FP32 steps, FP64 where they are
 * ill-conditioned). */
struct MixedLane {
    using Aim = double;
    RayConst<double> cd;
    RayConst<float> cf;
    RayState<float> s;
    Gathered<double> acc;
    double uHorizon, uEscape;
    float uNear, uEscapeF;
    Away<double> away;

    __device__ __forceinline__ void begin(const ArtfHipView& v, double nx, double ny) {
        Eye<double> e;
        eyeFrom64(v.frame64, e);
        RayState<double> sd;
        setupRay(e, nx, ny, cd, sd);
        cf = narrowConst(cd);
        s = castState<float>(sd);
        acc.r = acc.g = acc.b = 0.0, acc.through = 1.0;
        const float escape = v.escapeRadius > 0.f ? 1.f / v.escapeRadius : 0.f;
        uHorizon = e.uHorizon, uEscape = (double)escape;
        uNear = (float)(.9 * uHorizon), uEscapeF = escape;
    }

    __device__ __forceinline__ bool step(const ArtfHipView* views, int view, const float* map, float time, bool& sky) {
        RayState<float> n;
        StepLooks<float> looks;
        nystromStep(cf, s, n, looks);
        const bool special = s.w * n.w <= 0.f || s.v2 * n.v2 <= 0.f || s.n2 * n.n2 <= 0.f || n.u > uNear || s.u > uNear || !(n.u > uEscapeF);
        if (!special) {
            s = n;
            return false;
        }
        const RayState<double> w = castState<double>(s);
        RayState<double> m;
        StepLooks<double> l;
        nystromStep(cd, w, m, l);
        if (w.n2 * m.n2 < 0.0) {
            const ArtfHipView& v = views[view];
            DiscConst<double> d;
            discOfView(v, time, d);
            if (v.disc == DISC_PAINTED) crossPainted(d, cd, w, m, l, tablesOfView(v, map), acc);
            else crossAnalytic(d, cd, w, m, l, v.emissivity, map, acc);
        }
        const bool escaped = uEscape > m.u;
        if (escaped) {
            const double part = quot(w.u - uEscape, w.u - m.u);
            away.x = w.n0 + part * (m.n0 - w.n0);
            away.y = w.n1 + part * (m.n1 - w.n1);
            away.z = w.n2 + part * (m.n2 - w.n2);
            away.psi = w.psi + part * (m.psi - w.psi);
            sky = true;
        }
        if (escaped || m.u > uHorizon || !(acc.through > .004)) return true;
        s = castState<float>(m);
        return false;
    }

    __device__ __forceinline__ void finish(const ArtfHipView& v, const float* map, bool sky, float rgb[3]) {
        const float stars = v.frame.look[2];
        if (stars > 0.f && sky) skyLight(stars, cd, away.x, away.y, away.z, away.psi, (double)v.frame.look[1], map, acc);
        rgb[0] = (float)acc.r, rgb[1] = (float)acc.g, rgb[2] = (float)acc.b;
    }
};

template <int Precision> struct LaneOf { using type = PlainLane<float>; };
template <> struct LaneOf<ARTF_PRECISION_FP64> { using type = PlainLane<double>; };
template <> struct LaneOf<ARTF_PRECISION_MIXED> { using type = MixedLane; };

template <int Precision>
__global__ __launch_bounds__(kQueueGroup) void queueKernel(ArtfHipBatchArgs a) {
    using Lane = typename LaneOf<Precision>::type;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull; /* the lanes before this one */
    const int refillAt = a.schedule.refill, chunk = a.schedule.chunkTiles, stride = a.schedule.stepsPerCheck;

    Lane ray;
    bool active = false, pending = false, sky = false;
    int steps = 0, maxSteps = 0, view = 0;
    uint32_t outAt = 0;
    float time = 0.f;
    unsigned long long stepSum = 0, slots = 0;
    uint32_t poolNext = 0, poolEnd = 0; /* the wave's rays not yet given out (the same in every lane) */
    bool drained = false;

    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int idleCount = __popcll(idle);
        if (idleCount >= refillAt) {
            /* finish the rays that ended since the last refill, all together */
            if (pending) {
                const ArtfHipView& v = a.views[view];
                float rgb[3];
                ray.finish(v, a.map, sky, rgb);
                v.out[0][outAt] = rgb[0];
                v.out[1][outAt] = rgb[1];
                v.out[2][outAt] = rgb[2];
            }
            if (a.viewDone) {
                /* each view's count of rays done, one add per view among the finished lanes; the
                 * add that completes a view stamps the moment */
                unsigned long long left = __ballot(pending);
                while (left) {
                    const int first = __ffsll((long long)left) - 1;
                    const int v0 = __shfl(view, first);
                    const unsigned long long same = __ballot(pending && view == v0);
                    if (lane == first) {
                        const uint32_t count = (uint32_t)__popcll(same);
                        if (atomicAdd(&a.viewDone[v0], count) + count == a.views[v0].rays) a.stamps[1 + v0] = wall_clock64();
                    }
                    left &= ~same;
                }
            }
            pending = false;
            /* the pool's next rays to the idle lanes, in lane order */
            if (!drained) {
                const int rank = __popcll(idle & below);
                uint32_t mine = kNoRay;
                int given = 0;
                while (given < idleCount) {
                    if (poolNext == poolEnd) {
                        uint32_t base = 0;
                        if (lane == 0) base = atomicAdd(a.queue, (uint32_t)chunk);
                        base = __shfl(base, 0);
                        if (base >= a.tiles) {
                            drained = true;
                            break;
                        }
                        poolNext = base * 64u;
                        poolEnd = min(base + (uint32_t)chunk, a.tiles) * 64u;
                    }
                    const int take = min(idleCount - given, (int)(poolEnd - poolNext));
                    if (!active && rank >= given && rank < given + take) mine = poolNext + (uint32_t)(rank - given);
                    poolNext += (uint32_t)take;
                    given += take;
                }
                RayPlace p;
                if (mine != kNoRay && placeOf(a, mine, p)) {
                    const ArtfHipView& v = a.views[p.view];
                    const ArtfHoleFrame& f = v.frame;
                    const int side = (int)f.size[2];
                    float dx, dy;
                    subPixel(v.jitter, p.px, p.py, p.sub % side, p.sub / side, side, dx, dy);
                    time = rayMoment(f, p.sub, side);
                    typename Lane::Aim nx, ny;
                    aimOf(f, p.px, p.py, dx, dy, nx, ny);
                    ray.begin(v, nx, ny);
                    const uint32_t width = f.size[0], height = f.size[1];
                    outAt = (side == 1 ? 0u : (uint32_t)p.sub * width * height) + (uint32_t)p.py * width + (uint32_t)p.px;
                    view = p.view, maxSteps = v.maxSteps, steps = 0, sky = false, active = true;
                }
            }
            if (!__ballot(active)) {
                if (drained) break;
                continue;
            }
        }
        for (int k = 0; k < stride; ++k) {
            if (!__ballot(active)) break;
            slots += 64;
            if (active) {
                const bool ended = ray.step(a.views, view, a.map, time, sky);
                ++steps;
                if (ended || steps >= maxSteps) {
                    active = false, pending = true;
                    stepSum += (unsigned long long)steps;
                }
            }
        }
    }

    if (a.counters) {
        for (int offset = 32; offset > 0; offset >>= 1) stepSum += __shfl_xor(stepSum, offset);
        if (lane == 0) {
            const unsigned slot = (blockIdx.x * (kQueueGroup / 64) + (threadIdx.x >> 6)) % ARTF_HIP_COUNTER_SLOTS;
            atomicAdd(&a.counters[slot], stepSum);
            atomicAdd(&a.counters[ARTF_HIP_COUNTER_SLOTS + slot], slots);
        }
    }
}

template <int Precision> inline hipError_t launchQueueWith(const ArtfHipBatchArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(queueKernel<Precision>, dim3((unsigned)a.schedule.workgroups), dim3(kQueueGroup), 0, stream, a);
    return hipGetLastError();
}

template <int Precision> inline int queueOccupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, queueKernel<Precision>, kQueueGroup, 0) != hipSuccess) return 0;
    return blocks;
}

} // namespace artf
