/* The oracle's tracer, bloom and tone map (oracle.h). Test infrastructure: what the HIP path is
 * tested against, never a product path. Built by oracle/build.py with -O2 -ffp-contract=off -fno-math-errno
 * -fno-trapping-math, the reference's rules (CMakeLists.txt:54-58). */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* lens.inc:56-61: multiply-xor-shift on 32-bit words (the reference's lanes wrap as these do) */
static inline uint32_t oracleHash(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = (x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

/* float to int by truncation, as cvttps2dq for the values the tracer forms */
static inline int32_t oracleTrunc(double x) { return x != x ? INT32_MIN : (int32_t)x; }

#define REAL float
#define SUFFIX 32
#include "lens_ray.inc"
#undef REAL
#undef SUFFIX
#define REAL double
#define SUFFIX 64
#include "lens_ray.inc"
#undef REAL
#undef SUFFIX

static void eye32(const ArtfHoleFrame* f, Eye32* e) {
    for (int i = 0; i < 3; ++i) {
        e->forward[i] = f->forward[i], e->right[i] = f->right[i], e->up[i] = f->up[i];
        e->er[i] = f->er[i], e->etheta[i] = f->etheta[i], e->ephi[i] = f->ephi[i];
    }
    for (int i = 0; i < 4; ++i) e->metric[i] = f->metric[i], e->place[i] = f->place[i], e->boost[i] = f->boost[i];
    e->spin = f->spin[0], e->uHorizon = f->spin[1], e->discHeight = f->spin[2];
    e->inner = f->disc[0], e->outer = f->disc[1], e->thickness = f->disc[3];
}

static void eye64(const ArtfHoleFrame64* f, Eye64* e) {
    for (int i = 0; i < 3; ++i) {
        e->forward[i] = f->forward[i], e->right[i] = f->right[i], e->up[i] = f->up[i];
        e->er[i] = f->er[i], e->etheta[i] = f->etheta[i], e->ephi[i] = f->ephi[i];
    }
    for (int i = 0; i < 4; ++i) e->metric[i] = f->metric[i], e->place[i] = f->place[i], e->boost[i] = f->boost[i];
    e->spin = f->spin[0], e->uHorizon = f->spin[1], e->discHeight = f->spin[2];
    e->inner = f->disc[0], e->outer = f->disc[1], e->thickness = f->disc[3];
}

/* The place of ray (sx, sy) in pixel (px, py) (lens.inc:367-369), or a jittered place in the same
 * stratum (the same rule as the HIP path). */
static void placeOf(uint32_t seed, int px, int py, int sx, int sy, int side, float* dx, float* dy) {
    if (!seed) {
        *dx = ((float)sx + .5f) / (float)side;
        *dy = ((float)sy + .5f) / (float)side;
        return;
    }
    const uint32_t h = oracleHash((uint32_t)px ^ (seed * 0x9E3779B1u), (uint32_t)py, (uint32_t)(sy * side + sx) + seed);
    const float jx = ((float)(h & 0xFFFFu) + .5f) * (1.f / 65536.f), jy = ((float)(h >> 16) + .5f) * (1.f / 65536.f);
    *dx = ((float)sx + jx) / (float)side;
    *dy = ((float)sy + jy) / (float)side;
}

typedef struct TraceJob {
    const ArtfHoleFrame* f;
    const ArtfHoleFrame64* f64;
    const OracleLaw* law;
    int precision;
    const float* map;
    const float* fluid;
    float* out[3];
    uint32_t* steps;
    int rowBegin, rowEnd;
    int nextRow; /* under lock */
    pthread_mutex_t lock;
} TraceJob;

/* A row of pixels, each the mean of its side^2 rays in lensTile's order (lens.inc:361-380). */
static void traceRow(TraceJob* j, int y) {
    const ArtfHoleFrame* f = j->f;
    const int width = (int)f->size[0], height = (int)f->size[1], side = (int)f->size[2];
    Eye32 e32;
    Eye64 e64;
    eye32(f, &e32);
    if (j->precision == 64) eye64(j->f64, &e64);
    for (int x = 0; x < width; ++x) {
        float r = 0.f, g = 0.f, b = 0.f;
        double rd = 0, gd = 0, bd = 0;
        for (int sy = 0; sy < side; ++sy)
            for (int sx = 0; sx < side; ++sx) {
                float dx, dy;
                placeOf(j->law->jitter, x, y, sx, sy, side, &dx, &dy);
                const int sub = sy * side + sx;
                const float moment = ((float)(sub * 5 % (side * side)) + .5f) * f->image[3] - .5f;
                const float time = f->disc[2] + f->spin[3] * moment;
                int steps = 0;
                if (j->precision == 64) {
                    const double nx = ((double)x + (double)dx) * (2.0 / (double)width) - 1.0;
                    const double ny = 1.0 - ((double)y + (double)dy) * (2.0 / (double)height);
                    double tr, tg, tb;
                    ray64(f, &e64, j->law, j->map, j->fluid, nx, ny, time, &tr, &tg, &tb, &steps, NULL);
                    rd += tr, gd += tg, bd += tb;
                } else {
                    const float nx = ((float)(x & 7) + ((float)(x & ~7) + dx)) * f->image[1] - 1.f;
                    const float ny = 1.f - ((float)y + dy) * f->image[2];
                    float tr, tg, tb;
                    ray32(f, &e32, j->law, j->map, j->fluid, nx, ny, time, &tr, &tg, &tb, &steps, NULL);
                    r = r + tr, g = g + tg, b = b + tb;
                }
                if (j->steps) j->steps[(size_t)sub * width * height + (size_t)y * width + x] = (uint32_t)steps;
            }
        const size_t at = (size_t)y * width + x;
        if (j->precision == 64) {
            const double share = 1.0 / (double)(side * side);
            j->out[0][at] = (float)(rd * share), j->out[1][at] = (float)(gd * share), j->out[2][at] = (float)(bd * share);
        } else {
            j->out[0][at] = r * f->image[3], j->out[1][at] = g * f->image[3], j->out[2][at] = b * f->image[3];
        }
    }
}

static void* traceWorker(void* arg) {
    TraceJob* j = arg;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        const int y = j->nextRow++;
        pthread_mutex_unlock(&j->lock);
        if (y >= j->rowEnd) return NULL;
        traceRow(j, y);
    }
}

void oracleTrace(const ArtfHoleFrame* f, const ArtfHoleFrame64* f64, const OracleLaw* law, int precision, const float* map, const float* fluid, float* red,
                 float* green, float* blue, uint32_t* steps, int rowBegin, int rowEnd, int threads) {
    TraceJob j;
    memset(&j, 0, sizeof j);
    j.f = f, j.f64 = f64, j.law = law, j.precision = precision, j.map = map, j.fluid = fluid;
    j.out[0] = red, j.out[1] = green, j.out[2] = blue, j.steps = steps;
    j.rowBegin = rowBegin, j.rowEnd = rowEnd > 0 ? rowEnd : (int)f->size[1], j.nextRow = rowBegin;
    pthread_mutex_init(&j.lock, NULL);
    if (threads <= 0) {
        cpu_set_t set;
        threads = sched_getaffinity(0, sizeof set, &set) ? 1 : CPU_COUNT(&set);
    }
    if (threads > 64) threads = 64;
    pthread_t t[64];
    for (int i = 1; i < threads; ++i) pthread_create(&t[i], NULL, traceWorker, &j);
    traceWorker(&j);
    for (int i = 1; i < threads; ++i) pthread_join(t[i], NULL);
    pthread_mutex_destroy(&j.lock);
}

void oracleRay(const ArtfHoleFrame* f, const ArtfHoleFrame64* f64, const OracleLaw* law, int precision, const float* map, const float* fluid, double nx,
               double ny, float time, float* rgb, int* steps, double* path) {
    if (precision == 64) {
        Eye64 e;
        eye64(f64, &e);
        double r, g, b;
        ray64(f, &e, law, map, fluid, nx, ny, time, &r, &g, &b, steps, path);
        rgb[0] = (float)r, rgb[1] = (float)g, rgb[2] = (float)b;
    } else {
        Eye32 e;
        eye32(f, &e);
        float r, g, b;
        ray32(f, &e, law, map, fluid, (float)nx, (float)ny, time, &r, &g, &b, steps, path);
        rgb[0] = r, rgb[1] = g, rgb[2] = b;
    }
}

/* ---- the bloom pyramid (software.c:320-386, 419-426) ---- */

static inline float minF(float a, float b) { return b < a ? b : a; } /* artfMinf */
static inline float maxF(float a, float b) { return a < b ? b : a; } /* artfMaxf */
static inline int clampI(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static inline int minI(int a, int b) { return a < b ? a : b; }

static float tap(const float* level, int w, int h, float sx, float sy) {
    const float bx = floorf(sx), by = floorf(sy), px = sx - bx, py = sy - by;
    const int x0 = clampI((int)bx, 0, w - 1), x1 = clampI((int)bx + 1, 0, w - 1);
    const float* r0 = level + (size_t)clampI((int)by, 0, h - 1) * w;
    const float* r1 = level + (size_t)clampI((int)by + 1, 0, h - 1) * w;
    const float top = r0[x0] + (r0[x1] - r0[x0]) * px, bottom = r1[x0] + (r1[x1] - r1[x0]) * px;
    return top + (bottom - top) * py;
}

void oracleBloom(const ArtfHoleFrame* f, const float* const hdr[3], float* const bloom[3]) {
    const int width = (int)f->size[0], height = (int)f->size[1], levels = (int)f->size[3];
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < levels; ++k) {
            const uint32_t* to = f->bloom[k];
            for (uint32_t y = 0; y < to[2]; ++y) {
                float* out = bloom[c] + to[0] + (size_t)y * to[1];
                if (k == 0) {
                    const int y0 = minI((int)y * 2, height - 1), y1 = minI((int)y * 2 + 1, height - 1);
                    const float* a = hdr[c] + (size_t)y0 * width;
                    const float* b = hdr[c] + (size_t)y1 * width;
                    for (uint32_t x = 0; x < to[1]; ++x) {
                        const int x0 = minI((int)x * 2, width - 1), x1 = minI((int)x * 2 + 1, width - 1);
#define BRIGHT(v) minF(maxF((v) - 2.f, 0.f), 48.f)
                        out[x] = (BRIGHT(a[x0]) + BRIGHT(a[x1]) + BRIGHT(b[x0]) + BRIGHT(b[x1])) * .25f;
#undef BRIGHT
                    }
                } else {
                    const uint32_t* from = f->bloom[k - 1];
                    const float cy = (float)y * 2.f + .5f;
                    for (uint32_t x = 0; x < to[1]; ++x) {
                        const float cx = (float)x * 2.f + .5f;
                        float sum = 0;
                        for (int t = 0; t < 4; ++t)
                            sum += tap(bloom[c] + from[0], (int)from[1], (int)from[2], cx + (t & 1 ? 1.f : -1.f), cy + (t & 2 ? 1.f : -1.f));
                        out[x] = sum * .25f;
                    }
                }
            }
        }
        for (int k = levels - 2; k >= 0; --k) {
            const uint32_t* to = f->bloom[k];
            const uint32_t* from = f->bloom[k + 1];
            for (uint32_t y = 0; y < to[2]; ++y) {
                const float sy = ((float)y + .5f) * .5f - .5f;
                float* out = bloom[c] + to[0] + (size_t)y * to[1];
                for (uint32_t x = 0; x < to[1]; ++x) {
                    const float sx = ((float)x + .5f) * .5f - .5f;
                    float sum = 0;
                    for (int t = 0; t < 4; ++t)
                        sum += tap(bloom[c] + from[0], (int)from[1], (int)from[2], sx + (t & 1 ? .5f : -.5f), sy + (t & 2 ? .5f : -.5f));
                    out[x] += sum * (.25f * .85f);
                }
            }
        }
    }
}

/* ---- the tone map (lens.inc:387-440) ---- */

static float tone(float x) {
    const float top = x * (x * 2.51f + .03f);
    const float bottom = x * (x * 2.43f + .59f) + .14f;
    const float v = top / bottom;
    return sqrtf(v < 1.f ? v : 1.f);
}

void oracleToneMap(const ArtfHoleFrame* f, const float* const hdr[3], const float* const bloom[3], uint32_t* out) {
    const int width = (int)f->size[0], height = (int)f->size[1];
    const uint32_t* level = f->bloom[0];
    const int bw = (int)level[1], bh = (int)level[2];
    const float spread = f->look[3] * (1.f - .85f), lastX = (float)(bw - 1);
    for (int y = 0; y < height; ++y) {
        const float sy = ((float)y + .5f) * .5f - .5f, baseY = floorf(sy), partY = sy - baseY;
        const int y0 = clampI((int)baseY, 0, bh - 1), y1 = clampI((int)baseY + 1, 0, bh - 1);
        for (int x = 0; x < width; ++x) {
            const float sx = ((float)x + .5f) * .5f - .5f, baseX = floorf(sx), partX = sx - baseX;
            const float m0 = baseX > 0.f ? baseX : 0.f, m1 = baseX + 1.f > 0.f ? baseX + 1.f : 0.f;
            const int x0 = (int)(m0 < lastX ? m0 : lastX), x1 = (int)(m1 < lastX ? m1 : lastX);
            float glows[3];
            for (int c = 0; c < 3; ++c) {
                const float* p = bloom[c] + level[0];
                const float a = p[y0 * bw + x0], b = p[y0 * bw + x1], d = p[y1 * bw + x0], e = p[y1 * bw + x1];
                const float upper = a + (b - a) * partX, lower = d + (e - d) * partX;
                glows[c] = upper + (lower - upper) * partY;
            }
            const float grey = glows[0] * .3f + glows[1] * .59f + glows[2] * .11f;
            const uint32_t noise = oracleHash((uint32_t)(x + y * 8192), 7u, 13u);
            const float dither = (float)(int32_t)(noise & 255u) * (1.f / 255.f);
            uint32_t pixel = 0xFF000000u;
            for (int c = 0; c < 3; ++c) {
                const float glow = glows[c] + (grey - glows[c]) * .9f;
                const float light = hdr[c][(size_t)y * width + x] + glow * spread;
                const float v = tone(light * f->image[0]) * 255.f + dither;
                pixel |= (uint32_t)(int32_t)(v < 255.f ? v : 255.f) << (16 - 8 * c);
            }
            out[(size_t)y * width + x] = pixel;
        }
    }
}
