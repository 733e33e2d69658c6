/* artful-hip — the mixed-precision tracer: FP32 for the bulk of the Nystrom steps, FP64 where the
 * arithmetic is ill-conditioned.
 *
 * FP64 is kept for
 *   - the ray's setup: direction, aberration, the constants of motion (energy, xi, eta) and K, from
 *     the frame worked out in doubles; the FP32 coefficients are these, rounded once;
 *   - the radial turning point and the horizon: a step in which u' (w) changes sign, or whose FP32 u is
 *     past its capture value, has its radial half (u, u') taken again in FP64 from the FP32 state,
 *     widened; FP64 then decides whether the ray is captured there;
 *   - the polar turning point: a step in which m' (v.z) changes sign has its polar axis (m = n.z and m')
 *     taken again in FP64.
 * In Mino time u'' depends on u alone, and m'' = a^2 m - (K + 2 a^2 m^2) m on m alone, so what is redone
 * is the FP64 step's own result for those variables. The other two components of n (the azimuth) and
 * the drag angle psi, a quadrature of u, stay FP32, as does everything else: the disc-plane crossings,
 * the escape and the light they give are the FP32 arithmetic of tracer.h (as the FP32-with-FMA tracer
 * does it). Redoing the whole step in FP64 at each of these (drag included) gave the same error against
 * the FP64 tracer at 1080p (mean 2.33e-5 against 2.39e-5 now) and took 7% longer; taking the crossings and
 * the escape in FP64 as well gave a third of this mean error, but no more speed than the FP64 tracer
 * itself. Not in the reference (which is FP32 only). */
#pragma once

#include "tracer.h"

namespace artf {

template <typename To, typename From> __device__ __forceinline__ RayState<To> castState(const RayState<From>& s) {
    RayState<To> t;
    t.u = (To)s.u, t.w = (To)s.w, t.n0 = (To)s.n0, t.n1 = (To)s.n1, t.n2 = (To)s.n2;
    t.v0 = (To)s.v0, t.v1 = (To)s.v1, t.v2 = (To)s.v2, t.psi = (To)s.psi, t.drag = (To)s.drag;
    return t;
}

__device__ __forceinline__ RayConst<float> narrowConst(const RayConst<double>& c) {
    RayConst<float> t;
    t.a = (float)c.a, t.a2 = (float)c.a2, t.minusA = (float)c.minusA, t.twoA2 = (float)c.twoA2;
    t.big = (float)c.big, t.step = (float)c.step, t.step2 = (float)c.step2;
    t.uA = (float)c.uA, t.uB = (float)c.uB, t.uC = (float)c.uC, t.xiA = (float)c.xiA;
    t.xi = (float)c.xi, t.invEnergy = (float)c.invEnergy, t.moved = (float)c.moved;
    t.spinning = c.spinning;
    t.inRangeDrag = c.inRangeDrag;
    return t;
}

/* One mixed step: the FP32 step, then in FP64 again the radial half (u, u') where u' changes sign or u
 * passes the capture value, and the polar axis (m, m') where m' changes sign; the FP32 state takes the
 * FP64 results, rounded. */
__device__ __forceinline__ void mixedStep(const RayConst<double>& cd, const RayConst<float>& cf, const RayState<float>& s, float uHorizon,
                                          RayState<float>& n, StepLooks<float>& looks) {
    nystromStep(cf, s, n, looks);
    if (s.w * n.w <= 0.f || n.u > uHorizon) {
        double u, w, k1, midU;
        radialStep(cd, (double)s.u, (double)s.w, u, w, k1, midU);
        n.u = (float)u, n.w = (float)w, looks.k1 = (float)k1;
    }
    if (s.v2 * n.v2 <= 0.f) {
        double m, v, a1z;
        PolarFactors<double> k;
        polarAxisStep(cd, (double)s.n2, (double)s.v2, m, v, a1z, k);
        n.n2 = (float)m, n.v2 = (float)v, looks.a1z = (float)a1z;
    }
}

__device__ __forceinline__ RayResult traceRayMixed(const ArtfHoleFrame& f, const ArtfHoleFrame64& f64, const TraceLaw& law, const DiscTables& tables,
                                                   double nx, double ny, float time) {
    Eye<double> e;
    eyeFrom64(f64, e);
    RayConst<double> cd;
    RayState<double> sd;
    setupRay(e, nx, ny, cd, sd);
    RayConst<float> cf = narrowConst(cd);
    cf.inRangeDrag = law.inRangeDrag;
    RayState<float> s = castState<float>(sd);
    DiscConst<float> d;
    discOf<float>(f, f.disc[0], f.disc[1], f.disc[3], f.spin[2], time, d);
    Gathered<float> acc = {0.f, 0.f, 0.f, 1.f};
    bool sky = false;
    float awayX = 0, awayY = 0, awayZ = 0, awayPsi = 0;
    const float uHorizon = (float)e.uHorizon, uEscape = law.uEscape;
    int steps = law.maxSteps;
#pragma unroll 2 /* two steps a trip round the loop, as traceRay: 3.6% */
    for (int i = 0; i < law.maxSteps; ++i) {
        RayState<float> n;
        StepLooks<float> looks;
        mixedStep(cd, cf, s, uHorizon, n, looks);
        if (s.n2 * n.n2 < 0.f) {
            if (law.disc == DISC_PAINTED) crossPainted(d, cf, s, n, looks, tables, acc);
            else crossAnalytic(d, cf, s, n, looks, law.q, tables.map, acc);
        }
        const bool escaped = uEscape > n.u;
        if (escaped || n.u > uHorizon || !(acc.through > .004f)) { /* where it left, only as it leaves: as traceRay */
            if (escaped) {
                const float part = quot(s.u - uEscape, s.u - n.u);
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
    if (f.look[2] > 0.f && sky) skyLight(f.look[2], cf, awayX, awayY, awayZ, awayPsi, d.doppler, tables.map, acc);
    RayResult out;
    out.r = acc.r, out.g = acc.g, out.b = acc.b, out.steps = steps;
    return out;
}

} // namespace artf
