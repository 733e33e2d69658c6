/* artful-hip — what the headless drivers (artful-blackhole, artful-render) write: frames as images, and
 * each saved frame's scene as the bytes that made it. */
#ifndef ARTFUL_DRIVER_FRAMES_H
#define ARTFUL_DRIVER_FRAMES_H

#include <stdbool.h>
#include <stdio.h>

#include "artful/scene.h"

/* width x height BGRA8 into path: binary PPM (RGB) when the name ends in .ppm, else PNG (RGBA, 8 bits,
 * deflate's stored blocks: no compression, any reader takes it). false if it cannot be written. */
bool driverSaveImage(const char* path, const uint32_t* bgra, int width, int height);

/* One line: "frame <index> desc <hex of the ArtfBlackHoleDesc bytes> camera <hex of the ArtfCamera bytes>",
 * the description as given to artfSceneSetBlackHole (before its clamps) and the camera as set: what a
 * test needs to make the same frame again elsewhere. */
void driverDescribe(FILE* out, int index, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera);

#endif
