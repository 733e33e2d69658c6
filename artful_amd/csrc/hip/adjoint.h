/* artful-hip — the adjoint of the analytic-disc tracer: dL/d(parameters) for an image loss L, by
 * reverse accumulation along each ray with checkpoints.
 *
 * Forward, a ray is set up from the frame (direction, constants of motion E, xi, eta, K, the step h),
 * then stepped (tracer.h nystromStep) and, at each crossing of the disc's plane, shaded by the analytic
 * disc (tracer.h crossAnalytic): light += T alpha L(r) col(g), T *= 1 - alpha. Given dL/d(light) of the
 * ray (the pixel's seed, shared equally by its rays), the reverse pass walks the steps backwards:
 *   - the state (u, u', n, n', and the transmittance T) is kept every ARTF_ADJ_CHECKPOINT steps on the
 *     forward pass; each segment is recomputed from its checkpoint, its states held, then reversed
 *     (the last segment's states are held by the forward pass itself: a ray of at most
 *     ARTF_ADJ_CHECKPOINT steps is stepped forward once, not twice);
 *   - each step's adjoint is the transpose of its Nystrom update, with the step's stages recomputed;
 *   - each crossing's adjoint runs through the interpolated crossing point, the soft-edged window, the
 *     emissivity (inner/r)^q, the Doppler factor g^4 and the colour's heat g;
 *   - what the steps and crossings leave on the ray's constants runs back through its setup to the
 *     frame's fields (forward, right, up, metric, place, er, etheta, ephi), the spin, the disc's radii,
 *     q and the gain: ARTF_ADJ_COUNT numbers a ray, summed over rays by the kernel (adjoint.hip).
 * The host chains the frame's fields to the spin and the eye's position (hole.c artfHoleFrame64At).
 * Termination (capture, escape, opacity) and whether a step crosses are events: their derivatives are
 * zero almost everywhere and are taken as zero. Stars are off (the sky is not differentiable: hashed
 * cells), and the eye at rest (no aberration). Not in the reference. */
#pragma once

#include "tracer.h"

namespace artf {

#define ARTF_ADJ_CHECKPOINT 32 /* steps between checkpoints */
#define ARTF_ADJ_MAX_STEPS 2048
#define ARTF_ADJ_SLOTS (ARTF_ADJ_MAX_STEPS / ARTF_ADJ_CHECKPOINT)

/* The reverse pass's own arithmetic (the adjoints of a step, of a crossing and of the set-up) may fuse
 * a multiply with an add: it has only to be accurate, and a fused multiply-add rounds once where the two
 * operations rounded twice. The forward pass and its recomputation from the checkpoints stay unfused
 * (this file is built without contraction), so the states the adjoint reads are the tracer's, bit for
 * bit. Fused, the longest rays' reverse chains are shorter: the pass is 18% faster. */
#define ARTF_ADJ_FUSED _Pragma("clang fp contract(fast)")

/* Gradient layout (per ray, then summed): indices into ArtfHipAdjointArgs.partials' 32 words. */
enum {
    ADJ_FORWARD = 0, ADJ_RIGHT = 3, ADJ_UP = 6, ADJ_METRIC = 9, ADJ_PLACE = 13, ADJ_ER = 17, ADJ_ETHETA = 20, ADJ_EPHI = 23,
    ADJ_SPIN = 26, ADJ_INNER = 27, ADJ_OUTER = 28, ADJ_Q = 29, ADJ_GAIN = 30, ADJ_COUNT = 31
};

/* The adjoint of the ray constants the steps read. */
template <typename R> struct ConstBar {
    R step, big, twoA2, a2, uA, uB, uC;
};

/* One Nystrom step's adjoint: given the state s (the step is recomputed from it) and the adjoint nb
 * of the state after it, the adjoint of s (into sb), plus k1b, an extra adjoint on the step's first
 * look k1 = bend(u) (a crossing reads it). Accumulates the constants' adjoints into cb. */
template <typename R>
__device__ __forceinline__ void nystromStepAdjoint(const RayConst<R>& c, const RayState<R>& s, const RayState<R>& nb, R k1b, RayState<R>& sb,
                                                   ConstBar<R>& cb) {
    ARTF_ADJ_FUSED
    const R zero = 0;
    const R h = c.step, h2 = c.step2, big = c.big, t2 = c.twoA2, a2 = c.a2;
    const R half = h * (R).5f, eighth = h2 * (R).125f, halfSq = h2 * (R).5f, sixthSq = h2 * (R)(1.f / 6.f), sixth = h * (R)(1.f / 6.f);
    /* the stages, again */
    const R k1 = bendOf(c, s.u);
    const R c1 = big + t2 * (s.n2 * s.n2);
    const R a1[3] = {zero - c1 * s.n0, zero - c1 * s.n1, a2 * s.n2 - c1 * s.n2};
    const R midU = s.u + half * s.w + eighth * k1;
    const R m[3] = {s.n0 + half * s.v0 + eighth * a1[0], s.n1 + half * s.v1 + eighth * a1[1], s.n2 + half * s.v2 + eighth * a1[2]};
    const R k2 = bendOf(c, midU);
    const R c2 = big + t2 * (m[2] * m[2]);
    const R b1[3] = {zero - c2 * m[0], zero - c2 * m[1], a2 * m[2] - c2 * m[2]};
    const R endU = s.u + h * s.w + halfSq * k2;
    const R e[3] = {s.n0 + h * s.v0 + halfSq * b1[0], s.n1 + h * s.v1 + halfSq * b1[1], s.n2 + h * s.v2 + halfSq * b1[2]};
    const R k3 = bendOf(c, endU);
    const R c3 = big + t2 * (e[2] * e[2]);
    const R g1[3] = {zero - c3 * e[0], zero - c3 * e[1], a2 * e[2] - c3 * e[2]};
    const R n[3] = {s.n0, s.n1, s.n2}, v[3] = {s.v0, s.v1, s.v2};
    const R nnb[3] = {nb.n0, nb.n1, nb.n2}, nvb[3] = {nb.v0, nb.v1, nb.v2};

    R ub = zero, wb = zero, nbar[3] = {zero, zero, zero}, vbar[3] = {zero, zero, zero};
    R hb = zero, sixthSqb = zero, sixthb = zero, halfSqb = zero, halfb = zero, eighthb = zero;
    R k1bar = k1b, k2bar = zero, k3bar = zero;
    R a1b[3] = {zero, zero, zero}, b1b[3] = {zero, zero, zero}, g1b[3] = {zero, zero, zero};
    /* next u = u + h w + sixthSq (k1 + 2 k2); next w = w + sixth (k1 + k3 + 4 k2) */
    ub += nb.u;
    wb += h * nb.u;
    hb += s.w * nb.u;
    sixthSqb += (k1 + (k2 + k2)) * nb.u;
    k1bar += sixthSq * nb.u;
    k2bar += (R)2 * sixthSq * nb.u;
    wb += nb.w;
    sixthb += ((k1 + k3) + (R)4 * k2) * nb.w;
    k1bar += sixth * nb.w;
    k3bar += sixth * nb.w;
    k2bar += (R)4 * sixth * nb.w;
    for (int i = 0; i < 3; ++i) {
        nbar[i] += nnb[i];
        vbar[i] += h * nnb[i];
        hb += v[i] * nnb[i];
        sixthSqb += (a1[i] + (b1[i] + b1[i])) * nnb[i];
        a1b[i] += sixthSq * nnb[i];
        b1b[i] += (R)2 * sixthSq * nnb[i];
        vbar[i] += nvb[i];
        sixthb += ((a1[i] + g1[i]) + (R)4 * b1[i]) * nvb[i];
        a1b[i] += sixth * nvb[i];
        g1b[i] += sixth * nvb[i];
        b1b[i] += (R)4 * sixth * nvb[i];
    }
    /* the third look: g1 = -c3 e (+ a^2 e_z on z), c3 = K + 2a^2 e_z^2, k3 = bend(endU) */
    R eb[3] = {zero, zero, zero};
    const R c3b = -(e[0] * g1b[0] + e[1] * g1b[1] + e[2] * g1b[2]);
    eb[0] += -c3 * g1b[0];
    eb[1] += -c3 * g1b[1];
    eb[2] += (a2 - c3) * g1b[2];
    cb.a2 += e[2] * g1b[2];
    cb.big += c3b;
    cb.twoA2 += e[2] * e[2] * c3b;
    eb[2] += (R)2 * t2 * e[2] * c3b;
    const R endUb = (c.uA + endU * ((R)2 * c.uB + (R)3 * c.uC * endU)) * k3bar;
    cb.uA += endU * k3bar, cb.uB += endU * endU * k3bar, cb.uC += endU * endU * endU * k3bar;
    /* e = n + h v + halfSq b1; endU = u + h w + halfSq k2 */
    for (int i = 0; i < 3; ++i) {
        nbar[i] += eb[i];
        vbar[i] += h * eb[i];
        hb += v[i] * eb[i];
        halfSqb += b1[i] * eb[i];
        b1b[i] += halfSq * eb[i];
    }
    ub += endUb;
    wb += h * endUb;
    hb += s.w * endUb;
    halfSqb += k2 * endUb;
    k2bar += halfSq * endUb;
    /* the second look */
    R mb[3] = {zero, zero, zero};
    const R c2b = -(m[0] * b1b[0] + m[1] * b1b[1] + m[2] * b1b[2]);
    mb[0] += -c2 * b1b[0];
    mb[1] += -c2 * b1b[1];
    mb[2] += (a2 - c2) * b1b[2];
    cb.a2 += m[2] * b1b[2];
    cb.big += c2b;
    cb.twoA2 += m[2] * m[2] * c2b;
    mb[2] += (R)2 * t2 * m[2] * c2b;
    const R midUb = (c.uA + midU * ((R)2 * c.uB + (R)3 * c.uC * midU)) * k2bar;
    cb.uA += midU * k2bar, cb.uB += midU * midU * k2bar, cb.uC += midU * midU * midU * k2bar;
    /* m = n + half v + eighth a1; midU = u + half w + eighth k1 */
    for (int i = 0; i < 3; ++i) {
        nbar[i] += mb[i];
        vbar[i] += half * mb[i];
        halfb += v[i] * mb[i];
        eighthb += a1[i] * mb[i];
        a1b[i] += eighth * mb[i];
    }
    ub += midUb;
    wb += half * midUb;
    halfb += s.w * midUb;
    eighthb += k1 * midUb;
    k1bar += eighth * midUb;
    /* the first look */
    const R c1b = -(n[0] * a1b[0] + n[1] * a1b[1] + n[2] * a1b[2]);
    nbar[0] += -c1 * a1b[0];
    nbar[1] += -c1 * a1b[1];
    nbar[2] += (a2 - c1) * a1b[2];
    cb.a2 += n[2] * a1b[2];
    cb.big += c1b;
    cb.twoA2 += n[2] * n[2] * c1b;
    nbar[2] += (R)2 * t2 * n[2] * c1b;
    ub += (c.uA + s.u * ((R)2 * c.uB + (R)3 * c.uC * s.u)) * k1bar;
    cb.uA += s.u * k1bar, cb.uB += s.u * s.u * k1bar, cb.uC += s.u * s.u * s.u * k1bar;
    /* the step's coefficients: h/2, h^2/8, h^2/2, h^2/6, h/6 */
    cb.step += hb + halfb * (R).5f + sixthb * (R)(1.f / 6.f) + h * (eighthb * (R).25f + halfSqb + sixthSqb * (R)(1.f / 3.f));
    sb.u = ub, sb.w = wb;
    sb.n0 = nbar[0], sb.n1 = nbar[1], sb.n2 = nbar[2];
    sb.v0 = vbar[0], sb.v1 = vbar[1], sb.v2 = vbar[2];
    sb.psi = zero, sb.drag = zero;
}

/* What a crossing's adjoint leaves behind. */
template <typename R> struct CrossBar {
    R u, w, n2, nextN2, k1, step, a, invEnergy, xi, inner, outer, q, gain, through;
};

/* g = 1 + D (shift - 1) of tracer.h orbitShift, and its adjoint (gb: dL/dg). */
template <typename R>
__device__ __forceinline__ void orbitShiftAdjoint(const RayConst<R>& c, R uc, R doppler, R gb, R& ucb, R& ab, R& invEb, R& xib) {
    ARTF_ADJ_FUSED
    const R one = 1;
    const bool lowU = !(uc > (R)1e-4f);
    const R safeU = lowU ? (R)1e-4f : uc;
    const R rs = squareRoot(safeU);
    const R root = safeU * rs;
    const R lift = one + c.a * root;
    const R orbit = quot(root, lift);
    const R beneathRaw = one - (R)3 * safeU + ((R)2 * c.a) * root;
    const bool floorB = !(beneathRaw > (R)1e-4f);
    const R beneath = floorB ? (R)1e-4f : beneathRaw;
    const R sb = squareRoot(beneath);
    const R num = sb * (c.invEnergy * c.moved);
    const R den = lift * (one - orbit * c.xi);
    const R shift = quot(num, den);
    const R shiftb = doppler * gb;
    const R numb = quot(shiftb, den);
    const R denb = -shiftb * quot(shift, den);
    const R beneathb = floorB ? (R)0 : numb * quot(c.invEnergy * c.moved, (R)2 * sb);
    invEb += numb * sb * c.moved;
    R liftb = denb * (one - orbit * c.xi);
    const R orbitb = -denb * lift * c.xi;
    xib += -denb * lift * orbit;
    R rootb = quot(orbitb, lift);
    liftb += -orbitb * quot(root, lift * lift);
    R safeUb = (R)-3 * beneathb;
    ab += (R)2 * root * beneathb;
    rootb += (R)2 * c.a * beneathb;
    ab += root * liftb;
    rootb += c.a * liftb;
    safeUb += rootb * (R)1.5f * rs;
    ucb += lowU ? (R)0 : safeUb;
}

/* The adjoint of crossAnalytic (tracer.h) for one crossing, given the transmittance T before it, the
 * seed (dL/d of the ray's light, per channel) and the adjoint of T after it. */
template <typename R>
__device__ __forceinline__ bool crossAnalyticAdjoint(const DiscConst<R>& d, const RayConst<R>& c, const RayState<R>& s, const RayState<R>& n, R k1,
                                                     const float q, const float* __restrict__ map, R through, const R seed[3], R throughAfterb,
                                                     CrossBar<R>& xb) {
    ARTF_ADJ_FUSED
    const R zero = 0, one = 1;
    const R dd = s.n2 - n.n2;
    const R part = quot(s.n2, dd);
    const R hu = part * c.step, hh = hu * hu * (R).5f;
    const R uc = s.u + hu * s.w + hh * k1;
    if (!(uc > (R)1e-6f)) return false;
    const R r = quot(one, uc);
    const R edge = (R)ARTF_ANALYTIC_EDGE;
    const R x1 = quot(r - d.inner, edge) + (R).5f, x2 = quot(r - d.outer, edge) + (R).5f;
    const R w1 = edgeRamp(x1), w2 = edgeRamp(x2);
    const R window = w1 * (one - w2);
    if (!(window > zero)) return false;
    const R ratio = quot(d.inner, r);
    const R P = powPositive(ratio, (R)q);
    const R light0 = d.gain * P * window;
    R g = one, g4 = one;
    R heat = (R)(4500.f / (float)ARTF_HOLE_KELVIN);
    if (d.doppler != zero) {
        g = orbitShift(c, uc, d.doppler);
        const R g2 = g * g;
        g4 = g2 * g2;
        heat = heat * g;
    }
    const R light = light0 * g4;
    const float* colors = map + ARTF_HOLE_COLORS_AT;
    const R scaled = heat * (R)(float)(ARTF_HOLE_COLORS / 2);
    const R top = (R)(float)(ARTF_HOLE_COLORS - 1) - (R)1e-3f;
    const bool inRange = scaled > zero && scaled < top;
    const R at = lo(hi(scaled, zero), top);
    const R atFloor = floorOf(at), frac = at - atFloor;
    const int32_t c0 = toInt(atFloor) * 3;
    const R thickC = lo(hi(d.thickness, zero), one);
    const R alpha = thickC * window;
    const R weight = through * alpha * light;
    R col[3], slope[3];
    for (int k = 0; k < 3; ++k) {
        const R t0 = (R)colors[c0 + k], t1 = (R)colors[c0 + 3 + k];
        col[k] = t0 + (t1 - t0) * frac;
        slope[k] = (t1 - t0) * (R)(float)(ARTF_HOLE_COLORS / 2);
    }
    /* backwards */
    const R weightb = seed[0] * col[0] + seed[1] * col[1] + seed[2] * col[2];
    R heatb = inRange ? (seed[0] * weight * slope[0] + seed[1] * weight * slope[1] + seed[2] * weight * slope[2]) : zero;
    xb.through = throughAfterb * (one - alpha) + weightb * alpha * light;
    const R alphab = -throughAfterb * through + weightb * through * light;
    const R lightb = weightb * through * alpha;
    const R light0b = lightb * g4;
    R gb = zero;
    if (d.doppler != zero) gb = lightb * light0 * (R)4 * g * g * g + heatb * (R)(4500.f / (float)ARTF_HOLE_KELVIN);
    xb.gain = light0b * P * window;
    const R Pb = light0b * d.gain * window;
    const R windowb = light0b * d.gain * P + alphab * thickC;
    /* P = (inner / r)^q */
    xb.q = Pb * P * (R)log((double)ratio);
    R innerb = Pb * (R)q * quot(P, d.inner);
    R rb = -Pb * (R)q * quot(P, r);
    /* window = ramp(x1) (1 - ramp(x2)) */
    const R d1 = (x1 > zero && x1 < one) ? (R)6 * x1 * (one - x1) : zero;
    const R d2 = (x2 > zero && x2 < one) ? (R)6 * x2 * (one - x2) : zero;
    const R x1b = windowb * (one - w2) * d1, x2b = -windowb * w1 * d2;
    rb += quot(x1b + x2b, edge);
    innerb += -quot(x1b, edge);
    xb.inner = innerb;
    xb.outer = -quot(x2b, edge);
    R ucb = -rb * r * r;
    R ab = zero, invEb = zero, xib = zero;
    if (d.doppler != zero) orbitShiftAdjoint(c, uc, d.doppler, gb, ucb, ab, invEb, xib);
    xb.a = ab, xb.invEnergy = invEb, xb.xi = xib;
    /* uc = u + hu w + hh k1, hh = hu^2 / 2, hu = part h, part = n2 / (n2 - n2') */
    xb.u = ucb;
    xb.w = ucb * hu;
    const R hhb = ucb * k1;
    xb.k1 = ucb * hh;
    const R hub = ucb * s.w + hhb * hu;
    const R partb = hub * c.step;
    xb.step = hub * part;
    const R inv2 = quot(one, dd * dd);
    xb.n2 = partb * (-n.n2) * inv2;
    xb.nextN2 = partb * s.n2 * inv2;
    return true;
}

/* The per-ray gradient, and its setup adjoint: from the constants' and the initial state's adjoints
 * back to the frame's fields (tracer.h setupRay, with no aberration). */
template <typename R>
__device__ __forceinline__ void setupAdjoint(const Eye<R>& e, R nx, R ny, const RayConst<R>& c, const RayState<R>& s0b, const ConstBar<R>& cb,
                                             R invEb, R xib, R ab, R* grad) {
    ARTF_ADJ_FUSED
    const R zero = 0, one = 1;
    /* forward again: the direction and the constants */
    const R Dx = e.forward[0] + nx * e.right[0] + ny * e.up[0];
    const R Dy = e.forward[1] + nx * e.right[1] + ny * e.up[1];
    const R Dz = e.forward[2] + nx * e.right[2] + ny * e.up[2];
    const R len = squareRoot(Dx * Dx + Dy * Dy + Dz * Dz);
    const R invLength = quot(one, len);
    const R dx = Dx * invLength, dy = Dy * invLength, dz = Dz * invLength;
    const R a = c.a, a2 = c.a2, rho = e.place[0], rootDelta = e.place[1], u0 = e.place[2], mu0 = e.place[3];
    const R energy = e.metric[0] - e.metric[1] * dz;
    const R invE = quot(one, energy);
    const R xi = zero - e.metric[2] * dz * invE;
    const R xiAxis = e.metric[3] * dz * invE;
    const R across = rho * dy * invE;
    const R big = c.big;
    const R eta = big - xi * xi;
    /* constants: uA = a^2 - K, uB = 3 (eta + (xi - a)^2), uC = -2 a^2 eta, twoA2 = 2 a^2, h = S / sqrt(max(K, 4)) */
    R a2b = cb.a2 + cb.uA - (R)2 * eta * cb.uC + (R)2 * cb.twoA2;
    R bigb = cb.big - cb.uA;
    if (big > (R)4) bigb += cb.step * (-c.step * quot((R).5f, big));
    R etab = (R)3 * cb.uB - (R)2 * a2 * cb.uC;
    xib += (R)6 * (xi - a) * cb.uB;
    ab += (R)-6 * (xi - a) * cb.uB;
    etab += bigb;
    xib += (R)2 * xi * bigb;
    /* eta = across^2 + mu0^2 (xiAxis^2 - a^2) */
    R acrossb = (R)2 * across * etab;
    R mu0b = (R)2 * mu0 * (xiAxis * xiAxis - a2) * etab;
    R xiAxisb = (R)2 * mu0 * mu0 * xiAxis * etab;
    a2b += -mu0 * mu0 * etab;
    ab += (R)2 * a * a2b;
    /* initial state: u = u0, w = -u0^2 rootDelta rho dx invE, n = er, v = across etheta + xiAxis ephi */
    R placeb[4] = {zero, zero, s0b.u, mu0b};
    const R wb = s0b.w;
    placeb[2] += wb * ((R)-2 * u0 * rootDelta * rho * dx * invE);
    placeb[1] += wb * (-u0 * u0 * rho * dx * invE);
    placeb[0] += wb * (-u0 * u0 * rootDelta * dx * invE);
    R dxb = wb * (-u0 * u0 * rootDelta * rho * invE), dyb = zero, dzb = zero;
    invEb += wb * (-u0 * u0 * rootDelta * rho * dx);
    const R vb[3] = {s0b.v0, s0b.v1, s0b.v2};
    for (int i = 0; i < 3; ++i) {
        acrossb += e.etheta[i] * vb[i];
        xiAxisb += e.ephi[i] * vb[i];
        grad[ADJ_ETHETA + i] += across * vb[i];
        grad[ADJ_EPHI + i] += xiAxis * vb[i];
    }
    grad[ADJ_ER + 0] += s0b.n0, grad[ADJ_ER + 1] += s0b.n1, grad[ADJ_ER + 2] += s0b.n2;
    /* xi = -m2 dz invE, xiAxis = m3 dz invE, across = rho dy invE */
    R metricb[4] = {zero, zero, zero, zero};
    metricb[2] += -dz * invE * xib;
    dzb += -e.metric[2] * invE * xib;
    invEb += -e.metric[2] * dz * xib;
    metricb[3] += dz * invE * xiAxisb;
    dzb += e.metric[3] * invE * xiAxisb;
    invEb += e.metric[3] * dz * xiAxisb;
    placeb[0] += dy * invE * acrossb;
    dyb += rho * invE * acrossb;
    invEb += rho * dy * acrossb;
    /* invE = 1 / energy, energy = m0 - m1 dz */
    const R energyb = -invE * invE * invEb;
    metricb[0] += energyb;
    metricb[1] += -dz * energyb;
    dzb += -e.metric[1] * energyb;
    /* d = D / |D|, D = forward + nx right + ny up */
    const R dot = dx * dxb + dy * dyb + dz * dzb;
    const R Db[3] = {(dxb - dx * dot) * invLength, (dyb - dy * dot) * invLength, (dzb - dz * dot) * invLength};
    for (int i = 0; i < 3; ++i) {
        grad[ADJ_FORWARD + i] += Db[i];
        grad[ADJ_RIGHT + i] += nx * Db[i];
        grad[ADJ_UP + i] += ny * Db[i];
    }
    for (int i = 0; i < 4; ++i) grad[ADJ_METRIC + i] += metricb[i], grad[ADJ_PLACE + i] += placeb[i];
    grad[ADJ_SPIN] += ab;
}

/* A saved state: u, u', n, n' and the transmittance before the step's crossing. */
template <typename R> struct Saved {
    R u, w, n0, n1, n2, v0, v1, v2, through;
};

template <typename R> __device__ __forceinline__ Saved<R> save(const RayState<R>& s, R through) {
    Saved<R> k = {s.u, s.w, s.n0, s.n1, s.n2, s.v0, s.v1, s.v2, through};
    return k;
}

template <typename R> __device__ __forceinline__ RayState<R> restore(const Saved<R>& k) {
    RayState<R> s;
    s.u = k.u, s.w = k.w, s.n0 = k.n0, s.n1 = k.n1, s.n2 = k.n2, s.v0 = k.v0, s.v1 = k.v1, s.v2 = k.v2, s.psi = 0, s.drag = 0;
    return s;
}

/* One ray forward (its light into out[3]) and, given its seed, backwards: grad[ADJ_COUNT] += dL/d.
 * Returns the steps it took. ckpt: ARTF_ADJ_SLOTS saved states; seg: ARTF_ADJ_CHECKPOINT + 1 of them (a
 * segment's states and the one after its last step). */
template <typename R>
__device__ __forceinline__ int adjointRay(const ArtfHoleFrame& f, const Eye<R>& e, const TraceLaw& law, const float* __restrict__ map, R nx, R ny,
                                          float time, const R seed[3], float out[3], R* grad, Saved<R>* ckpt, Saved<R>* seg) {
    RayConst<R> c;
    RayState<R> s;
    setupRay(e, nx, ny, c, s);
    c.spinning = false; /* the drag angle shows only in the sky and the painted disc: not needed here */
    s.psi = 0, s.drag = 0;
    DiscConst<R> d;
    discFrom(f, e, time, d);
    Gathered<R> acc = {(R)0, (R)0, (R)0, (R)1};
    const R uHorizon = e.uHorizon, uEscape = (R)law.uEscape;
    const int maxSteps = min(law.maxSteps, ARTF_ADJ_MAX_STEPS);
    int steps = maxSteps;
    for (int i = 0; i < maxSteps; ++i) {
        if (i % ARTF_ADJ_CHECKPOINT == 0) ckpt[i / ARTF_ADJ_CHECKPOINT] = save(s, acc.through);
        seg[i % ARTF_ADJ_CHECKPOINT] = save(s, acc.through); /* what is left at the end: the last segment */
        RayState<R> n;
        StepLooks<R> looks;
        nystromStep(c, s, n, looks);
        if (s.n2 * n.n2 < (R)0) crossAnalytic(d, c, s, n, looks, law.q, map, acc);
        if (uEscape > n.u || n.u > uHorizon || !(acc.through > (R).004f)) {
            steps = i + 1;
            s = n;
            break;
        }
        s = n;
    }
    seg[(steps - 1) % ARTF_ADJ_CHECKPOINT + 1] = save(s, acc.through); /* the state after the last step */
    out[0] = (float)acc.r, out[1] = (float)acc.g, out[2] = (float)acc.b;

    /* backwards, a segment at a time */
    ConstBar<R> cb = {0, 0, 0, 0, 0, 0, 0};
    R invEb = 0, xib = 0, ab = 0;
    RayState<R> sb;
    sb.u = sb.w = sb.n0 = sb.n1 = sb.n2 = sb.v0 = sb.v1 = sb.v2 = sb.psi = sb.drag = (R)0;
    R throughb = 0;
    const int last = (steps - 1) / ARTF_ADJ_CHECKPOINT;
    for (int g = last; g >= 0; --g) {
        const int i0 = g * ARTF_ADJ_CHECKPOINT, i1 = min(i0 + ARTF_ADJ_CHECKPOINT, steps);
        if (g != last) { /* the segment's states again, from its checkpoint: the same bits as forward */
            RayState<R> t = restore(ckpt[g]);
            Gathered<R> tacc = {(R)0, (R)0, (R)0, ckpt[g].through};
            for (int i = i0; i < i1; ++i) {
                seg[i - i0] = save(t, tacc.through);
                RayState<R> n;
                StepLooks<R> looks;
                nystromStep(c, t, n, looks);
                if (t.n2 * n.n2 < (R)0) crossAnalytic(d, c, t, n, looks, law.q, map, tacc);
                t = n;
            }
            seg[i1 - i0] = save(t, tacc.through);
        }
        /* each step reads its state and the one after it from seg (the step is not taken again), the
         * next step's loaded a step ahead so that the load's wait overlaps this step's adjoint */
        Saved<R> after = seg[i1 - i0], here = seg[i1 - 1 - i0];
        for (int i = i1 - 1; i >= i0; --i) {
            const RayState<R> si = restore(here), n = restore(after);
            const R throughBefore = here.through;
            after = here;
            here = seg[max(i - 1 - i0, 0)];
            R k1b = 0, uDirect = 0, wDirect = 0, n2Direct = 0;
            if (si.n2 * n.n2 < (R)0) {
                CrossBar<R> xb;
                if (crossAnalyticAdjoint(d, c, si, n, bendOf(c, si.u), law.q, map, throughBefore, seed, throughb, xb)) {
                    sb.n2 += xb.nextN2;
                    k1b = xb.k1, uDirect = xb.u, wDirect = xb.w, n2Direct = xb.n2;
                    cb.step += xb.step;
                    ab += xb.a, invEb += xb.invEnergy, xib += xb.xi;
                    grad[ADJ_INNER] += xb.inner, grad[ADJ_OUTER] += xb.outer, grad[ADJ_Q] += xb.q, grad[ADJ_GAIN] += xb.gain;
                    throughb = xb.through;
                }
            }
            RayState<R> prev;
            nystromStepAdjoint(c, si, sb, k1b, prev, cb);
            prev.u += uDirect, prev.w += wDirect, prev.n2 += n2Direct;
            sb = prev;
        }
    }
    setupAdjoint(e, nx, ny, c, sb, cb, invEb, xib, ab, grad);
    return steps;
}

} // namespace artf
