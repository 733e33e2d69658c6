/* The oracle: a plain C11, scalar, CPU restatement of Artful 0.4.0's black-hole path, used only to
 * test the HIP path against (tests, the smoke check, the bench's CPU baseline). It is test
 * infrastructure and never a product path.
 *
 *   oracleTrace    src/software/lens.inc:99-383 (lensTrace + lensTile), FP32 bit for bit with the
 *                  reference's portable/AVX2 kernels (built -O3 -ffp-contract=off), or the same
 *                  algorithm in FP64 from an FP64 frame (the FP64 oracle), or with the analytic disc
 *   oracleBloom    src/software/software.c:320-386, 419-426
 *   oracleToneMap  src/software/lens.inc:387-440
 * Pinning (what each is checked against) is listed in the tests (tests/test_oracle.py). */
#ifndef ARTF_ORACLE_H
#define ARTF_ORACLE_H

#include <stdint.h>

#include "artful/hole.h"

typedef struct OracleLaw {
    int maxSteps;       /* 0: 192 */
    float escapeRadius; /* 0: where 1/r crosses 0 */
    int disc;           /* ArtfDiscModel */
    float emissivity;   /* q */
    uint32_t jitter;    /* 0: the stratified grid */
} OracleLaw;

/* Rows [rowBegin, rowEnd) of the HDR image into three planes of width*height floats. precision: 32
 * (from frame) or 64 (from frame64, which must then be given). steps: NULL, or per ray (side^2 per
 * pixel, ray k of pixel p at k*width*height + p): the Nystrom steps it took. threads: 0 all. */
void oracleTrace(const ArtfHoleFrame* frame, const ArtfHoleFrame64* frame64, const OracleLaw* law, int precision, const float* map, const float* fluid,
                 float* red, float* green, float* blue, uint32_t* steps, int rowBegin, int rowEnd, int threads);

/* One ray, for tests: its light, its steps, and (path: NULL or 6 doubles) where it left on the unit
 * sphere, its drag angle there, 1 if it left (0: captured, stopped or out of steps), and what of the
 * light behind got through. */
void oracleRay(const ArtfHoleFrame* frame, const ArtfHoleFrame64* frame64, const OracleLaw* law, int precision, const float* map, const float* fluid,
               double nx, double ny, float time, float* rgb, int* steps, double* path);

/* The bloom pyramid into three planes of the frame's bloomPixels each. */
void oracleBloom(const ArtfHoleFrame* frame, const float* const hdr[3], float* const bloom[3]);
/* BGRA8 of the HDR + bloom level 0. */
void oracleToneMap(const ArtfHoleFrame* frame, const float* const hdr[3], const float* const bloom[3], uint32_t* out);

#endif
