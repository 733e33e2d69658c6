/* artful-hip — the mixed-precision tracer: FP32 for the bulk of the Nystrom steps, FP64 where the
 * arithmetic is ill-conditioned.
 *
 * FP64 is kept for
 *   - the ray's setup: direction, aberration, the constants of motion (energy, xi, eta) and K, from
 *     the frame worked out in doubles; the FP32 coefficients are these, rounded once;
 *   - the radial and polar turning points: a step in which u' (w) or m' (v.z) changes sign is taken
 *     again in FP64 from the FP32 state, widened;
 *   - the disc-plane crossing and the escape, with their interpolation and the disc's light;
 *   - the steps near the horizon, where u passes .9 of its capture value.
 * The FP32 state after such a step is the FP64 result, rounded. Everything else is the FP32 step of
 * tracer.h. Not in the reference (which is FP32 only). */
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
    return t;
}

__device__ __forceinline__ RayResult traceRayMixed(const ArtfHoleFrame& f, const ArtfHoleFrame64& f64, const TraceLaw& law, const float* __restrict__ map,
                                   const float* __restrict__ fluid, double nx, double ny, float time) {
    Eye<double> e;
    eyeFrom64(f64, e);
    RayConst<double> cd;
    RayState<double> sd;
    setupRay(e, nx, ny, cd, sd);
    const RayConst<float> cf = narrowConst(cd);
    RayState<float> s = castState<float>(sd);
    DiscConst<double> d;
    discFrom(f, e, time, d);
    Gathered<double> acc = {0.0, 0.0, 0.0, 1.0};
    bool sky = false;
    double awayX = 0, awayY = 0, awayZ = 0, awayPsi = 0;
    const double uHorizon = e.uHorizon, uEscape = (double)law.uEscape;
    const float uNear = (float)(.9 * uHorizon), uEscapeF = law.uEscape;
    int steps = law.maxSteps;
    for (int i = 0; i < law.maxSteps; ++i) {
        RayState<float> n;
        StepLooks<float> looks;
        nystromStep(cf, s, n, looks);
        const bool special = s.w * n.w <= 0.f || s.v2 * n.v2 <= 0.f || s.n2 * n.n2 <= 0.f || n.u > uNear || s.u > uNear || !(n.u > uEscapeF);
        if (!special) {
            s = n;
            continue;
        }
        /* this step again, in FP64 */
        const RayState<double> w = castState<double>(s);
        RayState<double> m;
        StepLooks<double> l;
        nystromStep(cd, w, m, l);
        if (w.n2 * m.n2 < 0.0) {
            if (law.disc == DISC_PAINTED) crossPainted(d, cd, w, m, l, tablesOf(map, fluid), acc);
            else crossAnalytic(d, cd, w, m, l, law.q, map, acc);
        }
        const bool escaped = uEscape > m.u;
        if (escaped) {
            const double part = quot(w.u - uEscape, w.u - m.u);
            awayX = w.n0 + part * (m.n0 - w.n0);
            awayY = w.n1 + part * (m.n1 - w.n1);
            awayZ = w.n2 + part * (m.n2 - w.n2);
            awayPsi = w.psi + part * (m.psi - w.psi);
            sky = true;
        }
        if (escaped || m.u > uHorizon || !(acc.through > .004)) {
            steps = i + 1;
            break;
        }
        s = castState<float>(m);
    }
    if (f.look[2] > 0.f && sky) skyLight(f.look[2], cd, awayX, awayY, awayZ, awayPsi, d.doppler, map, acc);
    RayResult out;
    out.r = (float)acc.r, out.g = (float)acc.g, out.b = (float)acc.b, out.steps = steps;
    return out;
}

} // namespace artf
