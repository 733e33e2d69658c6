/* artful-hip — what the black-hole kernels read: the per-frame uniforms and the layout of the disc's
 * tables. Plain C, shared by the host code, the HIP kernels and the C-ABI.
 *
 * ArtfHoleFrame has the fields, order and 368-byte size of Artful 0.4.0's ArtfHoleFrame
 * (src/scene/sceneInternal.h:302-319); the map's layout is that of sceneInternal.h:17-33, 321-337,
 * so that the same bytes mean the same thing on both sides (golden vectors compare them directly). */
#ifndef ARTFUL_HOLE_H
#define ARTFUL_HOLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The painted gas: rings of it, a row each, rows even in 1/r. */
#define ARTF_HOLE_MAP_WIDTH 2048
#define ARTF_HOLE_MAP_HEIGHT 1024
#define ARTF_HOLE_DUST_WIDTH 1024
#define ARTF_HOLE_DUST_HEIGHT 256
#define ARTF_FLUID_WIDTH 1024 /* cells around a ring (a power of two) */
#define ARTF_FLUID_HEIGHT 256 /* rings */
#define ARTF_FLUID_DT .2f     /* the fluid's time a step, in masses */
#define ARTF_FLUID_WOUND 900.f /* how long its starting clumps have been sheared */
#define ARTF_FLUID_STEP_EDDIES 16 /* eddies stirred in a step: 2 large, then 14 small */
#define ARTF_HOLE_COLORS 256
#define ARTF_HOLE_SKY_WIDTH 1024
#define ARTF_HOLE_SKY_HEIGHT 512
#define ARTF_BLOOM_LEVELS 8
#define ARTF_BLOOM_CEILING 48.f   /* the most one pixel gives the bloom */
#define ARTF_BLOOM_FADE .85f      /* each level's weight against the one below */
#define ARTF_BLOOM_THRESHOLD 2.f  /* light past this feeds the bloom */
#define ARTF_BLOOM_WHITE .9f      /* how far the glow's colour goes to grey */

/* One block of floats: gas detail; per row (turns per unit time, heat); black-body colours; the
 * sky's dust; per row the light its gas gives; the disc's dust. */
#define ARTF_HOLE_ROWS_AT (ARTF_HOLE_MAP_WIDTH * ARTF_HOLE_MAP_HEIGHT)
#define ARTF_HOLE_COLORS_AT (ARTF_HOLE_ROWS_AT + ARTF_HOLE_MAP_HEIGHT * 2)
#define ARTF_HOLE_SKY_AT (ARTF_HOLE_COLORS_AT + ARTF_HOLE_COLORS * 3)
#define ARTF_HOLE_GLOW_AT (ARTF_HOLE_SKY_AT + ARTF_HOLE_SKY_WIDTH * ARTF_HOLE_SKY_HEIGHT)
#define ARTF_HOLE_DUST_AT (ARTF_HOLE_GLOW_AT + ARTF_HOLE_MAP_HEIGHT)
#define ARTF_HOLE_MAP_FLOATS (ARTF_HOLE_DUST_AT + ARTF_HOLE_DUST_WIDTH * ARTF_HOLE_DUST_HEIGHT)
#define ARTF_HOLE_FLUID .1f        /* the fluid's share of the gas's density */
#define ARTF_HOLE_DUST_DIM .97f    /* dust gives this much less light */
#define ARTF_HOLE_DUST_DENSE 25.f  /* and stops this much more */
#define ARTF_HOLE_KELVIN 6500.0    /* a heat of 1, in kelvin: the middle of the colour table */

/* How a ray is traced (artfRendererSetHoleTracing, artfHipTrace). */
typedef enum ArtfPrecision {
    ARTF_PRECISION_FP32 = 0,      /* the reference's FP32 arithmetic, operation for operation */
    ARTF_PRECISION_FP32_FAST = 1, /* FP32 with fused multiply-adds */
    ARTF_PRECISION_MIXED = 2,     /* FP64 where rays are set up, turn and end; FP32 steps between */
    ARTF_PRECISION_FP64 = 3       /* all of it in FP64 */
} ArtfPrecision;

typedef enum ArtfDiscModel {
    ARTF_DISC_PAINTED = 0,  /* the reference's disc: painted map x fluid x dust */
    ARTF_DISC_ANALYTIC = 1  /* a thin disc of emissivity (inner / r)^q, no map, fluid or dust */
} ArtfDiscModel;

/* The disc fluid's per-ring constants (Artful src/scene/fluid.c:255-285): radius, the span a ring's
 * cells cover in r, orbital rate, Kerr epicyclic swing, sound speed squared, cells a ring turns a step,
 * the envelope and falloff of what shows. Worked out on the host, the same for the CPU and GPU steps. */
typedef struct ArtfFluidRings {
    float r[ARTF_FLUID_HEIGHT], rowSpan[ARTF_FLUID_HEIGHT], omega[ARTF_FLUID_HEIGHT], swing[ARTF_FLUID_HEIGHT];
    float sound2[ARTF_FLUID_HEIGHT], shiftCells[ARTF_FLUID_HEIGHT], envelope[ARTF_FLUID_HEIGHT], falloff[ARTF_FLUID_HEIGHT];
} ArtfFluidRings;

/* One eddy stirred into the fluid (fluid.c:188-216): centred at radius r0 and cell i0, of the given
 * size, spinning with strength, adding clump to the gas; it touches cells within reach around. Its
 * parameters come from the fluid's random stream on the host; applying it is the same on CPU and GPU. */
typedef struct ArtfFluidEddy {
    float r0, size, strength, clump;
    int32_t i0, reach;
} ArtfFluidEddy;

/* Per-frame uniforms, in units of the hole's mass. */
typedef struct ArtfHoleFrame {
    float er[4];      /* the eye on the unit sphere (x, z, axis); w: its distance r */
    float etheta[4];  /* the two directions along the sphere there */
    float ephi[4];
    float right[4];   /* a pixel's ray = forward + nx right + ny up, in (r, theta, phi) at the eye */
    float up[4];
    float forward[4];
    float metric[4];  /* lapse, frame drag x varpi, varpi, varpi / sin(theta) */
    float place[4];   /* sqrt(Sigma), sqrt(Delta), 1/r, cos(theta) */
    float spin[4];    /* a, .995 / r_horizon, disc height, shutter */
    float flow[4];    /* the moment the fluid field is of */
    float boost[4];   /* the eye's velocity along r, theta, phi; w: gamma */
    float disc[4];    /* inner, outer, time, thickness */
    float look[4];    /* gain, doppler, stars, bloom */
    float image[4];   /* exposure, 2/width, 2/height, 1/(rays a pixel) */
    uint32_t size[4]; /* width, height, rays a side, bloom levels */
    uint32_t bloom[ARTF_BLOOM_LEVELS][4]; /* per level: first texel, width, height */
} ArtfHoleFrame;

/* The same geometry in double precision, worked out from the description in doubles: what the
 * FP64 and mixed tracers set each ray up from. */
typedef struct ArtfHoleFrame64 {
    double er[4], etheta[4], ephi[4];
    double right[4], up[4], forward[4];
    double metric[4], place[4], spin[4], boost[4];
    double disc[4];   /* inner, outer, time, thickness */
} ArtfHoleFrame64;

#ifdef __cplusplus
}
#endif

#endif
