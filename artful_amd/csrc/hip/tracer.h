/* artful-hip — one Kerr null geodesic per lane, for gfx950.
 *
 * The ray is traced backwards from the eye in Mino time with u = 1/r and n, a unit vector on the
 * sphere (so nothing is singular at the poles):
 *     u'' = u (A + u (B + u C)),  A = a^2 - K, B = 3 (eta + (xi - a)^2), C = -2 a^2 eta
 *     n'' = -(K + 2 a^2 m^2) n + a^2 m z^,  m = n.z
 *     psi' = -a u (2 - a xi u) / max(1 - 2u + a^2 u^2, 1e-3)          (frame dragging)
 * with K = eta + xi^2 from the constants of motion (energy, xi, eta) set up in the eye's ZAMO
 * frame; three-stage Nystrom steps of fixed size (pi/24)/sqrt(max(K,4)); the disc's plane crossed
 * where m changes sign; escape where u crosses 0 (or 1/escapeRadius); capture past .995/r_h; opaque
 * once the transmittance falls to .004. This is the algorithm of Artful 0.4.0's
 * src/software/lens.inc:99-351 (lensTrace) with its helpers lens.inc:30-96, restated per ray: the
 * reference runs 8 lanes in lockstep, but a lane's result depends on its own ray only (finished
 * lanes are masked), so one ray per GPU lane gives the same values.
 *
 * Real is float or double. With Real = float and contraction off (trace_exact.hip), every operation
 * is the reference's, in its association order, with IEEE division and square root and the
 * reference's min/max (a < b ? a : b, as x86's minps): the result is the reference's, bit for bit.
 * trace_fast.hip builds the same source with fused multiply-adds. Double gives the FP64 tracer. */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "artful/hole.h"

namespace artf {

/* The disc models and the tracer's step law. */
enum { DISC_PAINTED = 0, DISC_ANALYTIC = 1 };
#define ARTF_LENS_STEP .13089969f /* pi / 24 */
#define ARTF_LENS_SLICES 3
#define ARTF_TURN .15915494f      /* 1 / (2 pi) */
#define ARTF_ANALYTIC_EDGE 1.f    /* the analytic disc's soft edges, in masses */

template <typename R> __device__ __forceinline__ R lo(R a, R b) { return a < b ? a : b; }
template <typename R> __device__ __forceinline__ R hi(R a, R b) { return a > b ? a : b; }
__device__ __forceinline__ float squareRoot(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double squareRoot(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float floorOf(float x) { return floorf(x); }
__device__ __forceinline__ double floorOf(double x) { return floor(x); }
__device__ __forceinline__ float absOf(float x) { return fabsf(x); }
__device__ __forceinline__ double absOf(double x) { return fabs(x); }
/* x / y and the square root, correctly rounded: what hipcc emits for them by default
 * (-fhip-fp32-correctly-rounded-divide-sqrt); never the native one-ulp instructions. */
__device__ __forceinline__ float quot(float x, float y) { return x / y; }
__device__ __forceinline__ double quot(double x, double y) { return x / y; }
/* x / y correctly rounded by the steps of the IEEE sequence hipcc emits for it (the reciprocal refined once,
 * the quotient corrected twice), without its scaling (v_div_scale, twice) and fix-up (v_div_fixup): those
 * change nothing where y, 1/y and x/y are normal numbers and x is neither tiny nor zero, and here the sign
 * of a zero quotient is put back as IEEE gives it for y > 0 (x's). So, for 2^-20 <= y <= 2^20 and either
 * 2^-100 <= |x| <= 2^60, or x = 0, or y = 1 (the steps then return x itself, denormals included), the
 * quotient is x / y's, bit for bit, in 9 instructions instead of 11. Checked on the CPU with the hardware
 * reciprocal taken one ulp either way (tests/test_kernel_arith.py). */
__device__ __forceinline__ float quotInRange(float x, float y) {
    const float r0 = __builtin_amdgcn_rcpf(y);
    const float r = __builtin_fmaf(__builtin_fmaf(-y, r0, 1.f), r0, r0);
    const float q0 = x * r;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-y, q0, x), r, q0);
    return __builtin_copysignf(__builtin_fmaf(__builtin_fmaf(-y, q1, x), r, q1), x);
}
__device__ __forceinline__ double quotInRange(double x, double y) { return x / y; }
/* y + y + x, 4 y + x and the like with one rounding where the reference rounds twice. A product by 2 or 4
 * is exact short of overflow past 2^127, which no finite ray state comes near, so fusing it with the
 * addition rounds once, at the same place the addition did, to the same bits; zeros, denormals,
 * infinities and NaNs come out alike too (the product's sign is the operand's, and the sum of two zeros
 * follows one rule in both). A product by 1/2 or 1/4 is exact but in the denormal range, and where it is
 * used here (.5 m + .5, .25 warm + .75) a denormal-sized term lies far below the last bit of the sum, which
 * is then the addend either way. One instruction instead of two. */
__device__ __forceinline__ float scaledPlus(float k, float y, float x) { return __builtin_fmaf(k, y, x); }
__device__ __forceinline__ double scaledPlus(double k, double y, double x) { return __builtin_fma(k, y, x); }
/* float to int by truncation, as cvttps2dq for every value the tracer forms. */
template <typename R> __device__ __forceinline__ int32_t toInt(R x) { return (int32_t)x; }

/* The integer hash of lens.inc:56-61, on 32-bit words. */
__device__ __forceinline__ uint32_t lensHash(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = (x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

template <typename R> __device__ __forceinline__ R hashByte(uint32_t h, int shift) {
    return (R)(float)(int32_t)((h >> shift) & 255u) * (R)(1.f / 255.f);
}

/* atan2(y, x) in turns, by a minimax polynomial on [0, 1] (lens.inc:30-41). */
template <typename R> __device__ __forceinline__ R turnsOf(R y, R x) {
    const R ax = absOf(x), ay = absOf(y);
    const R most = hi(hi(ax, ay), (R)1e-20f), least = lo(ax, ay);
    const R t = quot(least, most), s = t * t;
    R p = (R)-.0464964749f * s + (R).15931422f;
    p = p * s - (R).327622764f;
    p = p * s * t + t;
    if (ay > ax) p = (R)1.57079633f - p;
    if (x < (R)0) p = (R)3.14159265f - p;
    if (y < (R)0) p = (R)0 - p;
    return p * (R)ARTF_TURN;
}

/* sin(2 pi t), folded to |t| <= 1/4 and a Taylor polynomial to x^9 (lens.inc:44-54). */
template <typename R> __device__ __forceinline__ R sineOfTurns(R turns) {
    R t = turns - floorOf(turns + (R).5f);
    if (t > (R).25f) t = (R).5f - t;
    if (t < (R)-.25f) t = (R)-.5f - t;
    const R x = t * (R)6.2831853f, x2 = x * x;
    R p = x2 * (R)2.7557319e-6f + (R)-1.9841270e-4f;
    p = p * x2 + (R)8.3333333e-3f;
    p = p * x2 + (R)-.16666667f;
    return p * x2 * x + x;
}

/* One layer of stars: hashed cubes of a lattice, a few holding a soft point (lens.inc:69-86). */
template <typename R>
__device__ __forceinline__ void starLayer(R dx_, R dy_, R dz_, float cells, int density, float size, float gain, R& r, R& g, R& b) {
    const R n = (R)cells;
    const R px = dx_ * n, py = dy_ * n, pz = dz_ * n;
    const R cx = floorOf(px), cy = floorOf(py), cz = floorOf(pz);
    const uint32_t h = lensHash((uint32_t)toInt(cx), (uint32_t)toInt(cy), (uint32_t)toInt(cz));
    if (!(density > (int32_t)(h & 255u))) return;
    const R ox = px - cx - hashByte<R>(h, 8), oy = py - cy - hashByte<R>(h, 16), oz = pz - cz - hashByte<R>(h, 24);
    const R d2 = ox * ox + oy * oy + oz * oz;
    R light = hi((R)1 - d2 * (R)(1.f / (size * size)), (R)0);
    light = light * light * (R)gain;
    const R warm = hashByte<R>(h, 3);
    r = r + light * scaledPlus((R).25f, warm, (R).75f); /* .75 + warm .25 */
    g = g + light * (R).82f;
    b = b + light * ((R)1 - warm * (R).45f);
}

/* A map row at a turn, linear between two columns that wrap (lens.inc:89-96). width: a power of 2. */
template <typename R> __device__ __forceinline__ R rowAt(const float* __restrict__ map, int width, int32_t row, R turn) {
    const R at = (turn - floorOf(turn)) * (R)width;
    const R column = floorOf(at), part = at - column;
    const int32_t base = row * width;
    const int32_t c0 = toInt(column) & (width - 1), c1 = (c0 + 1) & (width - 1);
    const R a = (R)map[base + c0], b = (R)map[base + c1];
    return a + (b - a) * part;
}

/* table[i] for i below 2^30, as a 32-bit byte offset from the table's base: the load takes it as it is. */
__device__ __forceinline__ float texel(const float* __restrict__ table, uint32_t i) {
    return *(const float*)((const char*)table + (i << 2));
}

/* rowAt for the disc's tables (gas, fluid, dust): their rows are never negative (rho in (0, 1) picks them,
 * where the sky's can fall to -1), so the offsets are unsigned 32-bit ones from the table's base, which the
 * loads take as they are (no 64-bit sign extension and add per load). The same texels, the same light. */
template <typename R> __device__ __forceinline__ R discRowAt(const float* __restrict__ table, uint32_t width, uint32_t row, R turn) {
    const R at = (turn - floorOf(turn)) * (R)(float)width;
    const R column = floorOf(at), part = at - column;
    const uint32_t base = row * width;
    const uint32_t c0 = (uint32_t)toInt(column) & (width - 1u), c1 = (c0 + 1u) & (width - 1u);
    const R a = (R)texel(table, base + c0), b = (R)texel(table, base + c1);
    return a + (b - a) * part;
}

/* What a ray needs of the frame, in its own precision. */
template <typename R> struct Eye {
    R forward[3], right[3], up[3];
    R er[3], etheta[3], ephi[3];
    R metric[4], place[4];
    R boost[4];
    R spin, uHorizon, discHeight, inner, outer, thickness;
};

template <typename R> __device__ __forceinline__ void eyeFrom(const ArtfHoleFrame& f, Eye<R>& e) {
    for (int i = 0; i < 3; ++i) {
        e.forward[i] = f.forward[i], e.right[i] = f.right[i], e.up[i] = f.up[i];
        e.er[i] = f.er[i], e.etheta[i] = f.etheta[i], e.ephi[i] = f.ephi[i];
    }
    for (int i = 0; i < 4; ++i) e.metric[i] = f.metric[i], e.place[i] = f.place[i], e.boost[i] = f.boost[i];
    e.spin = f.spin[0], e.uHorizon = f.spin[1], e.discHeight = f.spin[2];
    e.inner = f.disc[0], e.outer = f.disc[1], e.thickness = f.disc[3];
}

__device__ __forceinline__ void eyeFrom64(const ArtfHoleFrame64& f, Eye<double>& e) {
    for (int i = 0; i < 3; ++i) {
        e.forward[i] = f.forward[i], e.right[i] = f.right[i], e.up[i] = f.up[i];
        e.er[i] = f.er[i], e.etheta[i] = f.etheta[i], e.ephi[i] = f.ephi[i];
    }
    for (int i = 0; i < 4; ++i) e.metric[i] = f.metric[i], e.place[i] = f.place[i], e.boost[i] = f.boost[i];
    e.spin = f.spin[0], e.uHorizon = f.spin[1], e.discHeight = f.spin[2];
    e.inner = f.disc[0], e.outer = f.disc[1], e.thickness = f.disc[3];
}

/* The parts of the tracer's behaviour beyond the reference's. */
struct TraceLaw {
    int maxSteps;   /* the reference: 192 */
    float uEscape;  /* the reference: 0 */
    int disc;       /* DISC_PAINTED or DISC_ANALYTIC */
    float q;        /* the analytic disc's emissivity index */
    bool inRangeDrag; /* the steps' drag by quotInRange (RayConst::inRangeDrag); a kernel's constant */
};

struct RayResult {
    float r, g, b;
    int steps;
};

/* smoothstep on [0, 1] for the analytic disc's soft edges. */
template <typename R> __device__ __forceinline__ R edgeRamp(R t) {
    t = lo(hi(t, (R)0), (R)1);
    return t * t * scaledPlus((R)-2, t, (R)3); /* 3 - 2 t */
}

/* x^q for x > 0, through exp2/log2 in the ray's precision. */
__device__ __forceinline__ float powPositive(float x, float q) { return exp2f(q * log2f(x)); }
__device__ __forceinline__ double powPositive(double x, double q) { return exp2(q * log2(x)); }

/* What does not change along a ray: its constants of motion and the equations' coefficients. */
template <typename R> struct RayConst {
    R a, a2, minusA, twoA2;
    R big, step, step2;  /* K = eta + xi^2, the Mino step, its square */
    R uA, uB, uC, xiA;   /* u'' = u (uA + u (uB + u uC)); xi a */
    R xi, invEnergy, moved;
    bool spinning;
    bool inRangeDrag; /* the drag's quotient by quotInRange (the spin is 0 or at least 2^-50: dragOf) */
};

/* Where a ray is: u = 1/r and u', n on the unit sphere and n', the drag angle psi and its rate. */
template <typename R> struct RayState {
    R u, w, n0, n1, n2, v0, v1, v2, psi, drag;
};

/* What a disc crossing needs of the step it happened in. */
template <typename R> struct StepLooks {
    R k1, a1x, a1y, a1z;
};

/* The disc and the light it gives, per frame. */
template <typename R> struct DiscConst {
    R inner, outer, rhoScale, thickness, slant3, discHeight, gain, doppler, time, flowLag;
};

/* What a ray has gathered so far. */
template <typename R> struct Gathered {
    R r, g, b, through;
};

/* Where the painted disc's tables are: the map's shared part (gas, colours, sky, dust), and the ring
 * table (speed, heat per ring), glow per ring and fluid field of one view's disc. A scene's map holds
 * its own rings and glow (tablesOf); a batch of views keeps them apart, one set per view. */
struct DiscTables {
    const float* map;
    const float* rings;
    const float* glow;
    const float* fluid;
};

__device__ __forceinline__ DiscTables tablesOf(const float* map, const float* fluid) {
    DiscTables t;
    t.map = map, t.rings = map + ARTF_HOLE_ROWS_AT, t.glow = map + ARTF_HOLE_GLOW_AT, t.fluid = fluid;
    return t;
}

template <typename R> __device__ __forceinline__ R bendOf(const RayConst<R>& c, R x) { return x * (c.uA + x * (c.uB + x * c.uC)); }
/* The drag's rate -a x (2 - xi a x) / max(1 - 2x + a^2 x^2, 1e-3) at u = x. Its divisor lies in [1e-3, 2^20]
 * for |x| <= 2^9, and a stage point's u stays within a few units (the loop ends once u passes the capture
 * value, 1/r_h <= 1, and a step moves u by a fraction of itself). Where |x| < 2^-26 the divisor is 1 exactly
 * (1 - 2x rounds to 1 and a^2 x^2 is far below its last bit); elsewhere the dividend is 0 or at least |a| 2^-49
 * in size (2 - xi a x is 0 or a multiple of 2^-23, since xi a x near 2 leaves an exact difference, and
 * |a x| >= |a| 2^-26), so at least 2^-99 for a spin of at least 2^-50. quotInRange then gives the IEEE
 * quotient: c.inRangeDrag, set where the frame's spin is 0 or at least 2^-50 (artfHipDragInRange). */
template <typename R> __device__ __forceinline__ R dragOf(const RayConst<R>& c, R x) {
    const R num = c.minusA * x * ((R)2 - c.xiA * x), den = hi(scaledPlus((R)-2, x, (R)1) + c.a2 * x * x, (R)1e-3f); /* 1 - (x + x) */
    return c.inRangeDrag ? quotInRange(num, den) : quot(num, den);
}

/* The ray through a pixel, and its constants (lens.inc:101-163). */
template <typename R> __device__ __forceinline__ void setupRay(const Eye<R>& e, R nx, R ny, RayConst<R>& c, RayState<R>& s) {
    const R zero = 0, one = 1;
    R dx = e.forward[0] + nx * e.right[0] + ny * e.up[0];
    R dy = e.forward[1] + nx * e.right[1] + ny * e.up[1];
    R dz = e.forward[2] + nx * e.right[2] + ny * e.up[2];
    const R invLength = quot(one, squareRoot(dx * dx + dy * dy + dz * dz));
    dx = dx * invLength, dy = dy * invLength, dz = dz * invLength;

    /* a moving eye: aberration toward its motion, and the Doppler factor it brings */
    R moved = one;
    const R gamma = e.boost[3];
    if (gamma > one) {
        const R bx = e.boost[0], by = e.boost[1], bz = e.boost[2], share = quot(gamma, gamma + one);
        const R cc = zero - (dx * bx + dy * by + dz * bz);
        const R push = gamma * (one + share * cc);
        const R over = quot(one, gamma * (one + cc));
        dx = (dx - bx * push) * over;
        dy = (dy - by * push) * over;
        dz = (dz - bz * push) * over;
        moved = over;
    }

    /* constants of motion, from the direction in the eye's ZAMO frame */
    const R a = e.spin, a2 = a * a, rho = e.place[0], rootDelta = e.place[1], u0 = e.place[2], mu0 = e.place[3];
    const R energy = e.metric[0] - e.metric[1] * dz;
    const R invEnergy = quot(one, energy);
    const R xi = zero - e.metric[2] * dz * invEnergy;
    const R xiAxis = e.metric[3] * dz * invEnergy;
    const R across = rho * dy * invEnergy;
    const R eta = across * across + (mu0 * mu0) * (xiAxis * xiAxis - a2);
    const R big = eta + xi * xi;
    c.a = a, c.a2 = a2, c.minusA = -a, c.twoA2 = (R)2 * a2;
    c.big = big;
    c.step = quot((R)ARTF_LENS_STEP, squareRoot(hi(big, (R)4)));
    c.step2 = c.step * c.step;
    c.uA = a2 - big;
    const R xiOff = xi - a;
    c.uB = (R)3 * (eta + xiOff * xiOff);
    c.uC = ((R)-2 * a2) * eta;
    c.xiA = xi * a;
    c.xi = xi, c.invEnergy = invEnergy, c.moved = moved;
    c.spinning = a != zero;
    c.inRangeDrag = false; /* the caller's to set, for the steps */

    s.u = u0;
    s.w = ((-u0 * u0 * rootDelta * rho) * dx) * invEnergy;
    s.n0 = e.er[0], s.n1 = e.er[1], s.n2 = e.er[2];
    s.v0 = across * e.etheta[0] + xiAxis * e.ephi[0];
    s.v1 = across * e.etheta[1] + xiAxis * e.ephi[1];
    s.v2 = across * e.etheta[2] + xiAxis * e.ephi[2];
    s.psi = zero;
    s.drag = c.spinning ? dragOf(c, s.u) : zero;
}

/* The radial half of a Nystrom step (lens.inc:167-192): u and u' = w, which depend on u alone; also the
 * first stage's u'' (k1) and the middle stage's u (midU), which the crossing and the drag read. */
template <typename R>
__device__ __forceinline__ void radialStep(const RayConst<R>& c, R u, R w, R& nu, R& nw, R& k1, R& midU) {
    const R step = c.step, step2 = c.step2;
    const R half = step * (R).5f, eighth = step2 * (R).125f, halfSquare = step2 * (R).5f;
    const R sixthSquare = step2 * (R)(1.f / 6.f), sixth = step * (R)(1.f / 6.f);
    k1 = bendOf(c, u);
    midU = u + half * w + eighth * k1;
    const R k2 = bendOf(c, midU);
    const R endU = u + step * w + halfSquare * k2;
    const R k3 = bendOf(c, endU);
    /* k1 + (k2 + k2) and (k1 + k3) + 4 k2, each with its exact doubling or quadrupling fused */
    nu = u + step * w + sixthSquare * scaledPlus((R)2, k2, k1);
    nw = w + sixth * scaledPlus((R)4, k2, k1 + k3);
}

/* The three stages' factors K + 2 a^2 m^2 of the polar equations. */
template <typename R> struct PolarFactors {
    R c1, c2, c3;
};

/* The polar half's axis component: m = n.z and m' = v.z, whose equation m'' = a^2 m - (K + 2 a^2 m^2) m
 * involves m alone; also its first stage's m'' and the stages' factors, which the other two components
 * take. */
template <typename R>
__device__ __forceinline__ void polarAxisStep(const RayConst<R>& c, R m, R v, R& nm, R& nv, R& a1z, PolarFactors<R>& k) {
    const R step = c.step, step2 = c.step2, big = c.big, twoA2 = c.twoA2, a2 = c.a2;
    const R half = step * (R).5f, eighth = step2 * (R).125f, halfSquare = step2 * (R).5f;
    const R sixthSquare = step2 * (R)(1.f / 6.f), sixth = step * (R)(1.f / 6.f);
    k.c1 = big + twoA2 * (m * m);
    a1z = a2 * m - k.c1 * m;
    const R m2 = m + half * v + eighth * a1z;
    k.c2 = big + twoA2 * (m2 * m2);
    const R b1z = a2 * m2 - k.c2 * m2;
    const R e2 = m + step * v + halfSquare * b1z;
    k.c3 = big + twoA2 * (e2 * e2);
    const R g1z = a2 * e2 - k.c3 * e2;
    /* a + (b + b) and (a + g) + 4 b, each with its exact doubling or quadrupling fused */
    nm = m + step * v + sixthSquare * scaledPlus((R)2, b1z, a1z);
    nv = v + sixth * scaledPlus((R)4, b1z, a1z + g1z);
}

/* The polar half: n on the unit sphere and n', which depend on n alone; also the first stage's n''. */
template <typename R>
__device__ __forceinline__ void polarStep(const RayConst<R>& c, const RayState<R>& s, RayState<R>& n, StepLooks<R>& looks) {
    const R zero = 0;
    const R step = c.step, step2 = c.step2;
    const R half = step * (R).5f, eighth = step2 * (R).125f, halfSquare = step2 * (R).5f;
    const R sixthSquare = step2 * (R)(1.f / 6.f), sixth = step * (R)(1.f / 6.f);
    PolarFactors<R> k;
    polarAxisStep(c, s.n2, s.v2, n.n2, n.v2, looks.a1z, k);
    const R a1x = zero - k.c1 * s.n0, a1y = zero - k.c1 * s.n1;
    const R m0 = s.n0 + half * s.v0 + eighth * a1x, m1 = s.n1 + half * s.v1 + eighth * a1y;
    const R b1x = zero - k.c2 * m0, b1y = zero - k.c2 * m1;
    const R e0 = s.n0 + step * s.v0 + halfSquare * b1x, e1 = s.n1 + step * s.v1 + halfSquare * b1y;
    const R g1x = zero - k.c3 * e0, g1y = zero - k.c3 * e1;
    n.n0 = s.n0 + step * s.v0 + sixthSquare * scaledPlus((R)2, b1x, a1x);
    n.n1 = s.n1 + step * s.v1 + sixthSquare * scaledPlus((R)2, b1y, a1y);
    n.v0 = s.v0 + sixth * scaledPlus((R)4, b1x, a1x + g1x);
    n.v1 = s.v1 + sixth * scaledPlus((R)4, b1y, a1y + g1y);
    looks.a1x = a1x, looks.a1y = a1y;
}

/* One Nystrom step (lens.inc:167-192): the radial and polar halves (the equations for u and for n do not
 * touch each other in Mino time), then the drag angle psi by Simpson's rule, from u alone. */
template <typename R>
__device__ __forceinline__ void nystromStep(const RayConst<R>& c, const RayState<R>& s, RayState<R>& n, StepLooks<R>& looks) {
    const R zero = 0;
    R midU;
    radialStep(c, s.u, s.w, n.u, n.w, looks.k1, midU);
    polarStep(c, s, n, looks);
    n.drag = zero, n.psi = zero;
    if (c.spinning) {
        const R sixth = c.step * (R)(1.f / 6.f);
        n.drag = dragOf(c, n.u);
        n.psi = s.psi + sixth * scaledPlus((R)4, dragOf(c, midU), s.drag + n.drag);
    }
}

/* The Doppler and gravitational shift g of gas on a circular orbit at uc, for this ray
 * (lens.inc:269-275). */
template <typename R> __device__ __forceinline__ R orbitShift(const RayConst<R>& c, R uc, R doppler) {
    const R one = 1;
    const R safeU = hi(uc, (R)1e-4f);
    const R root = safeU * squareRoot(safeU);
    const R lift = one + c.a * root;
    const R orbit = quot(root, lift);
    const R beneath = hi(one - (R)3 * safeU + ((R)2 * c.a) * root, (R)1e-4f);
    const R shift = quot(squareRoot(beneath) * (c.invEnergy * c.moved), lift * (one - orbit * c.xi));
    return one + doppler * (shift - one);
}

/* The painted disc where the ray crosses its plane in this step (lens.inc:195-291): three slices
 * through the disc's depth, front to back, then black-body colour at the shifted heat. */
template <typename R>
__device__ __forceinline__ void crossPainted(const DiscConst<R>& d, const RayConst<R>& c, const RayState<R>& s, const RayState<R>& n,
                                             const StepLooks<R>& l, const DiscTables& tables, Gathered<R>& acc) {
    const R zero = 0, one = 1;
    const R part = quot(s.n2, s.n2 - n.n2);
    const R hu = part * c.step, hh = hu * hu * (R).5f;
    const R uc = s.u + hu * s.w + hh * l.k1;
    const R rhoCenter = (one - d.inner * uc) * d.rhoScale;
    if (!(rhoCenter > (R)-.03f && rhoCenter < (R)1.02f)) return;
    const float* __restrict__ map = tables.map;
    const float* __restrict__ rows = tables.rings;
    const float* __restrict__ colors = map + ARTF_HOLE_COLORS_AT;
    const R wc = s.w + hu * l.k1;
    const R cx = s.n0 + hu * s.v0 + hh * l.a1x, cy = s.n1 + hu * s.v1 + hh * l.a1y;
    const R vx = s.v0 + hu * l.a1x, vy = s.v1 + hu * l.a1y, vz = s.v2 + hu * l.a1z;
    const R psiC = s.psi + part * (n.psi - s.psi);
    const R dragC = s.drag + part * (n.drag - s.drag);
    const R turn = turnsOf(cy, cx) + psiC * (R)ARTF_TURN;
    const R rate = (quot(cx * vy - cy * vx, hi(cx * cx + cy * cy, (R)1e-6f)) + dragC) * (R)ARTF_TURN;
    /* gas seen at a slant is deeper */
    const R slope = quot(wc, uc);
    const R speed2 = (vx * vx + vy * vy) + (vz * vz + slope * slope);
    const R mu = hi(quot(absOf(vz), squareRoot(speed2)), (R).02f);
    const R slant = quot(d.slant3, mu);
    const R reach = lo(quot(d.discHeight, hi(absOf(vz), (R)1e-5f)), c.step * (R)1.5f);
    R glow = zero, clear = acc.through;
    for (int k = 0; k < ARTF_LENS_SLICES; ++k) {
        const R t = (R)(((float)k + .5f) * (2.f / (float)ARTF_LENS_SLICES) - 1.f);
        const R ds = reach * t;
        const R uk = uc + wc * ds;
        const R rhoK = (one - d.inner * uk) * d.rhoScale;
        if (!(rhoK > zero && rhoK < one)) continue; /* outside the disc: adds nothing */
        const R rowAtF = rhoK * (R)(float)(ARTF_HOLE_MAP_HEIGHT - 1);
        const R rowFloor = lo(floorOf(rowAtF), (R)(float)(ARTF_HOLE_MAP_HEIGHT - 2));
        const R rowPart = rowAtF - rowFloor;
        const uint32_t row0 = (uint32_t)toInt(rowFloor), row1 = row0 + 1u; /* rhoK in (0, 1): 0..1023 */
        const R speed0 = (R)texel(rows, row0 + row0), speed1 = (R)texel(rows, row1 + row1);
        const R turnK = turn + rate * ds;
        const R d0 = discRowAt(map, ARTF_HOLE_MAP_WIDTH, row0, turnK - speed0 * d.time);
        const R d1 = discRowAt(map, ARTF_HOLE_MAP_WIDTH, row1, turnK - speed1 * d.time);
        R density = d0 + (d1 - d0) * rowPart;
        /* the fluid as last stepped, each ring turned on since */
        const R fluidAt = rhoK * (R)(float)(ARTF_FLUID_HEIGHT - 1);
        const R fluidFloor = lo(floorOf(fluidAt), (R)(float)(ARTF_FLUID_HEIGHT - 2));
        const uint32_t fluidRow = (uint32_t)toInt(fluidFloor);
        const R fluidTurn = turnK - speed0 * d.flowLag;
        const R q0 = discRowAt(tables.fluid, ARTF_FLUID_WIDTH, fluidRow, fluidTurn), q1 = discRowAt(tables.fluid, ARTF_FLUID_WIDTH, fluidRow + 1u, fluidTurn);
        const R gas = q0 + (q1 - q0) * (fluidAt - fluidFloor);
        density = density * ((R)(1.f - ARTF_HOLE_FLUID) + (R)ARTF_HOLE_FLUID * gas);
        const R dustAt = rhoK * (R)(float)(ARTF_HOLE_DUST_HEIGHT - 1);
        const uint32_t dustRow = (uint32_t)toInt(lo(floorOf(dustAt), (R)(float)(ARTF_HOLE_DUST_HEIGHT - 1)));
        const R dusty = discRowAt(map + ARTF_HOLE_DUST_AT, ARTF_HOLE_DUST_WIDTH, dustRow, turnK - speed0 * d.time);
        const R lit = one - (R)ARTF_HOLE_DUST_DIM * dusty;
        const R height = t * quot(reach, d.discHeight);
        const R level = hi(one - vz * vz * (height * height), zero);
        density = density * level;
        const R depth = slant * density * (one + (R)ARTF_HOLE_DUST_DENSE * dusty);
        const R alpha = quot(depth, one + depth);
        glow = glow + clear * alpha * (density * lit);
        clear = clear * (one - alpha);
    }
    /* temperature, and what the gas's orbit and the hole's pull do to it */
    const R rhoSafe = lo(hi(rhoCenter, zero), one);
    const uint32_t rowIndex = (uint32_t)toInt(rhoSafe * (R)(float)(ARTF_HOLE_MAP_HEIGHT - 1)); /* rhoSafe in [0, 1] */
    R heat = (R)texel(rows, rowIndex + rowIndex + 1u);
    const R gives = (R)texel(tables.glow, rowIndex);
    R light = d.gain * glow * gives;
    if (d.doppler != zero) {
        const R g = orbitShift(c, uc, d.doppler);
        const R g2 = g * g;
        light = light * (g2 * g2);
        heat = heat * g;
    }
    const int32_t color = toInt(lo(heat * (R)(float)(ARTF_HOLE_COLORS / 2), (R)(float)(ARTF_HOLE_COLORS - 1)));
    const int32_t at = color + color + color;
    acc.r = acc.r + light * (R)colors[at];
    acc.g = acc.g + light * (R)colors[at + 1];
    acc.b = acc.b + light * (R)colors[at + 2];
    acc.through = clear;
}

/* A thin disc of emissivity (inner/r)^q with soft edges ARTF_ANALYTIC_EDGE wide, opacity
 * min(thickness, 1), Doppler-boosted as the painted one, its colour the black-body table read
 * linearly. Not in the reference: the smooth disc the precision study and the adjoint use. */
template <typename R>
__device__ __forceinline__ void crossAnalytic(const DiscConst<R>& d, const RayConst<R>& c, const RayState<R>& s, const RayState<R>& n,
                                              const StepLooks<R>& l, const float q, const float* __restrict__ map, Gathered<R>& acc) {
    const R zero = 0, one = 1;
    const R part = quot(s.n2, s.n2 - n.n2);
    const R hu = part * c.step, hh = hu * hu * (R).5f;
    const R uc = s.u + hu * s.w + hh * l.k1;
    if (!(uc > (R)1e-6f)) return;
    const R r = quot(one, uc);
    const R edge = (R)ARTF_ANALYTIC_EDGE;
    const R window = edgeRamp(quot(r - d.inner, edge) + (R).5f) * (one - edgeRamp(quot(r - d.outer, edge) + (R).5f));
    if (!(window > zero)) return;
    R light = d.gain * powPositive(quot(d.inner, r), (R)q) * window;
    R heat = (R)(4500.f / (float)ARTF_HOLE_KELVIN);
    if (d.doppler != zero) {
        const R g = orbitShift(c, uc, d.doppler);
        const R g2 = g * g;
        light = light * (g2 * g2);
        heat = heat * g;
    }
    const float* colors = map + ARTF_HOLE_COLORS_AT;
    const R at = lo(hi(heat * (R)(float)(ARTF_HOLE_COLORS / 2), zero), (R)(float)(ARTF_HOLE_COLORS - 1) - (R)1e-3f);
    const R atFloor = floorOf(at), frac = at - atFloor;
    const int32_t c0 = toInt(atFloor) * 3;
    const R alpha = lo(hi(d.thickness, zero), one) * window;
    const R weight = acc.through * alpha * light;
    acc.r = acc.r + weight * ((R)colors[c0] + ((R)colors[c0 + 3] - (R)colors[c0]) * frac);
    acc.g = acc.g + weight * ((R)colors[c0 + 1] + ((R)colors[c0 + 4] - (R)colors[c0 + 1]) * frac);
    acc.b = acc.b + weight * ((R)colors[c0 + 2] + ((R)colors[c0 + 5] - (R)colors[c0 + 2]) * frac);
    acc.through = acc.through * (one - alpha);
}

/* Stars and the sky's dust toward where the ray left, rotated by its drag (lens.inc:316-347). */
template <typename R>
__device__ __forceinline__ void skyLight(float stars, const RayConst<R>& c, R awayX, R awayY, R awayZ, R awayPsi, R doppler,
                                         const float* __restrict__ map, Gathered<R>& acc) {
    const R zero = 0, one = 1;
    const R invAway = quot(one, squareRoot(hi(awayX * awayX + awayY * awayY + awayZ * awayZ, (R)1e-12f)));
    const R psiTurns = awayPsi * (R)ARTF_TURN;
    const R sine = sineOfTurns(psiTurns), cosine = sineOfTurns(psiTurns + (R).25f);
    const R x = awayX * invAway, y = awayY * invAway, m = awayZ * invAway;
    const R sx = x * cosine - y * sine, sy = x * sine + y * cosine;
    R r = zero, g = zero, b = zero;
    starLayer(sx, sy, m, 90.f, 26, .1f, 1.6f, r, g, b);
    starLayer(sx, sy, m, 240.f, 70, .17f, .22f, r, g, b);
    const R skyTurn = turnsOf(sy, sx);
    const R heightAt = scaledPlus((R).5f, m, (R).5f) * (R)(float)(ARTF_HOLE_SKY_HEIGHT - 1); /* m .5 + .5 */
    const R skyFloor = lo(floorOf(heightAt), (R)(float)(ARTF_HOLE_SKY_HEIGHT - 2));
    const int32_t skyRow = toInt(skyFloor);
    const R s0 = rowAt(map + ARTF_HOLE_SKY_AT, ARTF_HOLE_SKY_WIDTH, skyRow, skyTurn);
    const R s1 = rowAt(map + ARTF_HOLE_SKY_AT, ARTF_HOLE_SKY_WIDTH, skyRow + 1, skyTurn);
    const R dust = s0 + (s1 - s0) * (heightAt - skyFloor);
    r = r + dust * (R).92f;
    g = g + dust * (R).86f;
    b = b + dust;
    R shown = acc.through * (R)stars;
    if (doppler != zero) {
        const R g1 = one + doppler * (c.invEnergy * c.moved - one);
        const R g2 = g1 * g1;
        shown = shown * (g2 * g2);
    }
    acc.r = acc.r + r * shown;
    acc.g = acc.g + g * shown;
    acc.b = acc.b + b * shown;
}

/* The disc's constants for a ray at this moment; inner, outer, thickness and height in R (from the
 * FP32 frame, or the FP64 one). */
template <typename R>
__device__ __forceinline__ void discOf(const ArtfHoleFrame& f, R inner, R outer, R thickness, R discHeight, float time, DiscConst<R>& d) {
    d.inner = inner, d.outer = outer, d.thickness = thickness, d.discHeight = discHeight;
    d.rhoScale = quot((R)1, (R)1 - quot(inner, outer));
    d.slant3 = (R)(f.disc[3] / (float)ARTF_LENS_SLICES);
    d.gain = (R)f.look[0], d.doppler = (R)f.look[1];
    d.time = (R)time, d.flowLag = (R)(time - f.flow[0]);
}

template <typename R> __device__ __forceinline__ void discFrom(const ArtfHoleFrame& f, const Eye<R>& e, float time, DiscConst<R>& d) {
    discOf(f, e.inner, e.outer, e.thickness, e.discHeight, time, d);
}

/* One ray, eye to end, all in R. nx, ny: the pixel's place in -1..1; time: this ray's moment. */
template <typename R>
__device__ __forceinline__ RayResult traceRay(const ArtfHoleFrame& f, const Eye<R>& e, const TraceLaw& law, const DiscTables& tables, R nx, R ny,
                                              float time) {
    RayConst<R> c;
    RayState<R> s;
    setupRay(e, nx, ny, c, s);
    c.inRangeDrag = law.inRangeDrag;
    DiscConst<R> d;
    discFrom(f, e, time, d);
    Gathered<R> acc = {(R)0, (R)0, (R)0, (R)1};
    bool sky = false;
    R awayX = 0, awayY = 0, awayZ = 0, awayPsi = 0;
    const R uHorizon = e.uHorizon, uEscape = (R)law.uEscape;
    int steps = law.maxSteps;
#pragma unroll 2 /* two steps a trip round the loop: fewer branches and moves, 2.4% (5% with FMA) */
    for (int i = 0; i < law.maxSteps; ++i) {
        RayState<R> n;
        StepLooks<R> looks;
        nystromStep(c, s, n, looks);
        if (s.n2 * n.n2 < (R)0) { /* through the disc's plane */
            if (law.disc == DISC_PAINTED) crossPainted(d, c, s, n, looks, tables, acc);
            else crossAnalytic(d, c, s, n, looks, law.q, tables.map, acc);
        }
        const bool escaped = uEscape > n.u;
        /* A lane leaves the loop at escape, at the horizon or opaque. Where an escaping ray left is worked
         * out here, where only waves with a lane leaving go: as a branch of its own every step it cost every
         * wave its exec-mask work at every step, and after the loop it kept both states alive across it
         * (18 more register moves a step). */
        if (escaped || n.u > uHorizon || !(acc.through > (R).004f)) {
            if (escaped) {
                const R part = quot(s.u - uEscape, s.u - n.u);
                awayX = s.n0 + part * (n.n0 - s.n0);
                awayY = s.n1 + part * (n.n1 - s.n1);
                awayZ = s.n2 + part * (n.n2 - s.n2);
                awayPsi = s.psi + part * (n.psi - s.psi);
                sky = true;
            }
            steps = i + 1;
            break;
        }
        s = n;
    }
    if (f.look[2] > 0.f && sky) skyLight(f.look[2], c, awayX, awayY, awayZ, awayPsi, d.doppler, tables.map, acc);
    RayResult out;
    out.r = (float)acc.r, out.g = (float)acc.g, out.b = (float)acc.b, out.steps = steps;
    return out;
}

/* A frame's ray: the rings and glow inside its map (the frame's map layout). */
template <typename R>
__device__ __forceinline__ RayResult traceRay(const ArtfHoleFrame& f, const Eye<R>& e, const TraceLaw& law, const float* __restrict__ map,
                                              const float* __restrict__ fluid, R nx, R ny, float time) {
    return traceRay(f, e, law, tablesOf(map, fluid), nx, ny, time);
}

} // namespace artf
