/* artful-hip — the black hole's disc data made on the GPU (beyond the reference, which builds its map
 * and steps its fluid on the CPU for every backend, src/scene/hole.c:136-197, src/scene/fluid.c): the
 * same map and the same fluid, bit for bit, left in device memory for the tracer. The host keeps only
 * the fluid's random stream, its clock and the small tables (colours, ring constants). */
#ifndef ARTFUL_DEVICE_DISC_H
#define ARTFUL_DEVICE_DISC_H

#include "artful/scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A disc map (ARTF_HOLE_MAP_FLOATS floats, the layout of artful/hole.h) in device memory on HIP device
 * `device`: the gas, the disc's dust and the sky painted there (buildDisc, hole.c:136-178), the colour
 * table from the host; its per-ring rows and glow zero until artfDeviceMapRings. NULL if it cannot be
 * made there (logged). Free with artfDeviceMapFree. */
float* artfDeviceMapNew(int device);
/* The per-ring rows (speed, heat) and glow of this disc (setRings, hole.c:186-197) into a device map,
 * once the device's queued work (a renderer's trace reading the map) is done. */
bool artfDeviceMapRings(float* map, float inner, float outer, float spin, int device);
/* Frees it once the device's queued work is done. */
void artfDeviceMapFree(float* map);

typedef struct ArtfDeviceFluid ArtfDeviceFluid;

/* As artfFluidNew(inner, outer, spin, time), on HIP device `device`; NULL if it cannot run there (logged). */
ArtfDeviceFluid* artfDeviceFluidNew(float inner, float outer, float spin, double time, int device);
/* Frees it once the device's queued work (a renderer's trace reading the field) is done. */
void artfDeviceFluidFree(ArtfDeviceFluid* f);
/* As artfFluidAdvance: true if the field changed. The new field is ready when it returns. */
bool artfDeviceFluidAdvance(ArtfDeviceFluid* f, double time);
/* The same, queued on the fluid's stream without waiting: the new field goes to the other of two field
 * buffers, after every read of that buffer marked with artfDeviceFluidReadDone has ended; a reader makes
 * its stream wait for it with artfDeviceFluidReadOn. */
bool artfDeviceFluidAdvanceQueued(ArtfDeviceFluid* f, double time);
/* The field in device memory, ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT floats: ready when New or Advance
 * returns, or after AdvanceQueued on a stream that artfDeviceFluidReadOn has ordered. */
const float* artfDeviceFluidField(const ArtfDeviceFluid* f);
/* Work queued on `stream` from now on waits for the current field to be written. */
bool artfDeviceFluidReadOn(ArtfDeviceFluid* f, void* stream);
/* The work queued on `stream` so far reads the current field: the fluid writes that buffer again only
 * after it. */
bool artfDeviceFluidReadDone(ArtfDeviceFluid* f, void* stream);
/* The field copied to the host. */
bool artfDeviceFluidRead(const ArtfDeviceFluid* f, float* field);
double artfDeviceFluidTime(const ArtfDeviceFluid* f);

/* Where a scene keeps its black hole's disc data: device >= 0 builds its map (artfDeviceMapNew) and
 * steps its fluid (ArtfDeviceFluid) on that HIP device — same bits, same calls: artfSceneSetBlackHole
 * makes, re-rings or advances them — and a HIP renderer on the same device reads them in place, with
 * no host work beyond the small tables and no upload; -1 (the default) keeps them on the host, as the
 * reference does. A scene that already has a disc builds it again in the new place (its gas starts
 * again at the description's time). Falls back to the host, logged, if the device cannot hold them. */
void artfSceneSetDiscOnDevice(ArtfScene* scene, int device);

#ifdef __cplusplus
}
#endif

#endif
