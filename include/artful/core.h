/* artful-hip — core types and services of the black-hole path: memory, the log, the clock, and the
 * small vector type the scene's API is written in.
 *
 * Restates the names and signatures of Artful 0.4.0's include/artful/core.h:28-35 (memory),
 * :55-60 (log, clock), :73-94 (numbers, vectors), so that code written against Artful compiles here.
 * A pointer a function returns is the caller's only where the function says so. */
#ifndef ARTFUL_CORE_H
#define ARTFUL_CORE_H

#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARTF_VERSION_MAJOR 0
#define ARTF_VERSION_MINOR 4
#define ARTF_VERSION_PATCH 0

#define ARTF_INLINE static inline

/* Running out of host memory aborts (core.h:28): callers never check for NULL. */
void* artfMalloc(size_t size);
void* artfCalloc(size_t count, size_t size);
void* artfRealloc(void* p, size_t size);
void artfFree(void* p);
void* artfAlignedMalloc(size_t size, size_t align);
void artfAlignedFree(void* p);

#define ARTF_NEW(T) ((T*)artfCalloc(1, sizeof(T)))

/* Where messages go: nothing is said until a function is given; any thread may log. */
typedef void (*ArtfLogFn)(void* user, const char* message);
void artfSetLog(ArtfLogFn fn, void* user);
void artfLog(const char* format, ...);

/* A monotonic clock in milliseconds. */
double artfNowMs(void);

ARTF_INLINE float artfMinf(float a, float b) { return b < a ? b : a; }
ARTF_INLINE float artfMaxf(float a, float b) { return a < b ? b : a; }
ARTF_INLINE float artfClampf(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }
ARTF_INLINE int artfMini(int a, int b) { return a < b ? a : b; }
ARTF_INLINE int artfMaxi(int a, int b) { return a > b ? a : b; }
ARTF_INLINE int artfClampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

typedef struct ArtfVec3 { float x, y, z; } ArtfVec3;

ARTF_INLINE ArtfVec3 artfVec3(float x, float y, float z) { ArtfVec3 v = {x, y, z}; return v; }
ARTF_INLINE ArtfVec3 artfVec3Add(ArtfVec3 a, ArtfVec3 b) { return artfVec3(a.x + b.x, a.y + b.y, a.z + b.z); }
ARTF_INLINE ArtfVec3 artfVec3Sub(ArtfVec3 a, ArtfVec3 b) { return artfVec3(a.x - b.x, a.y - b.y, a.z - b.z); }
ARTF_INLINE ArtfVec3 artfVec3Scale(ArtfVec3 a, float s) { return artfVec3(a.x * s, a.y * s, a.z * s); }
/* (summed left to right: the frame's setup depends on these exact roundings) */
ARTF_INLINE float artfVec3Dot(ArtfVec3 a, ArtfVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ARTF_INLINE ArtfVec3 artfVec3Cross(ArtfVec3 a, ArtfVec3 b) {
    return artfVec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
ARTF_INLINE float artfVec3Length(ArtfVec3 a) { return sqrtf(artfVec3Dot(a, a)); }
/* The zero vector stays zero. */
ARTF_INLINE ArtfVec3 artfVec3Normalize(ArtfVec3 a) {
    const float l = artfVec3Length(a);
    return l > 0 ? artfVec3Scale(a, 1.f / l) : a;
}

#ifdef __cplusplus
}
#endif

#endif
