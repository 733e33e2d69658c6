/* artful-hip — the FP64 frame of a hole (frame64.c). Declared over the public types alone
 * (artful/scene.h: ArtfBlackHoleDesc, ArtfCamera), which this library and the reference lay out alike. */
#ifndef ARTFUL_FRAME64_H
#define ARTFUL_FRAME64_H

#include "artful/scene.h"

struct ArtfHoleFrame64;

/* d: clamped (artfSceneSetBlackHole keeps it so). */
void artfHoleFrame64Of(const ArtfBlackHoleDesc* d, const ArtfCamera* c, int width, int height, struct ArtfHoleFrame64* out);
/* The same as a function of a spin and an eye position given in doubles (for derivatives). */
void artfHoleFrame64At(const ArtfBlackHoleDesc* d, const ArtfCamera* c, double spin, const double eye[3], int width, int height,
                       struct ArtfHoleFrame64* out);

#endif
