/* artful-hip — arithmetic that the host C (scene/disc.c, fluid.c) and the HIP kernels (hip/disc.hip,
 * fluid.hip) share, written once: plain C11 for gcc, host-and-device functions under hipcc. Both sides
 * build it without contraction and with IEEE division, so the same source gives the same bits on either.
 * The clamp, min and max forms are the reference's (artfClampf, artfMinf, artfMaxf: which operand is
 * kept when they compare equal, and NaN's way through, matter to a chaotic fluid). */
#ifndef ARTFUL_SHARED_MATH_H
#define ARTFUL_SHARED_MATH_H

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define ARTF_SHARED __host__ __device__ static inline
#else
#define ARTF_SHARED static inline
#endif

ARTF_SHARED float sharedClamp(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }
ARTF_SHARED float sharedMin(float a, float b) { return b < a ? b : a; }
ARTF_SHARED float sharedMax(float a, float b) { return a < b ? b : a; }

/* Smoothstep of x between lo and hi. */
ARTF_SHARED float sharedSmoothStep(float lo, float hi, float x) {
    const float t = sharedClamp((x - lo) / (hi - lo), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}

/* The lattice hash of the reference's value noise (hole.c:14-21, fluid.c:60-66): multiply-xor-shift. */
ARTF_SHARED uint32_t sharedLattice(uint32_t x, uint32_t y, uint32_t seed) {
    uint32_t h = (x * 0x9E3779B1u) ^ (y * 0x85EBCA77u) ^ (seed * 0xC2B2AE3Du);
    h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    h = (h ^ (h >> 12)) * 0x297A2D39u;
    return h ^ (h >> 15);
}

#endif
