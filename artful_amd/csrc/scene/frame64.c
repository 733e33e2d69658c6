/* artful-hip — a frame's geometry worked out in doubles (what the FP64 and mixed tracers set rays up
 * from), from a hole's clamped description and a camera: the body of the reference's float frame
 * (src/scene/hole.c:273-356) in double precision. It reads nothing but its arguments, so it links into
 * this library and into a build of the reference's tree with the HIP backend alike. */
#include "frame64.h"

#include <math.h>
#include <string.h>

#include "artful/hole.h"
#include "frame_body.h"

#define CAMERA_EYE(k) c->position.k

void artfHoleFrame64Of(const ArtfBlackHoleDesc* d, const ArtfCamera* c, int width, int height, ArtfHoleFrame64* out) {
    memset(out, 0, sizeof *out);
    HOLE_FRAME_BODY(double, sqrt, tan, fabs, (double)d->spin, CAMERA_EYE)
}

/* The FP64 frame as a function of a spin and an eye position given in doubles (for derivatives). */
void artfHoleFrame64At(const ArtfBlackHoleDesc* d, const ArtfCamera* c, double spin, const double eye[3], int width, int height, ArtfHoleFrame64* out) {
    memset(out, 0, sizeof *out);
#define GIVEN_EYE(k) eye[#k[0] - 'x']
    HOLE_FRAME_BODY(double, sqrt, tan, fabs, spin, GIVEN_EYE)
#undef GIVEN_EYE
}

