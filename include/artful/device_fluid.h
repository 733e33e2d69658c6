/* artful-hip — the disc fluid stepped on the GPU (beyond the reference, which steps it on the CPU for
 * every backend, src/scene/fluid.c): the same fluid as artfFluidNew / artfFluidAdvance
 * (artful/inspect.h), bit for bit, its field left in device memory for the tracer. The host keeps
 * only the random stream, the clock and the ring constants. */
#ifndef ARTFUL_DEVICE_FLUID_H
#define ARTFUL_DEVICE_FLUID_H

#include "artful/core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ArtfDeviceFluid ArtfDeviceFluid;

/* As artfFluidNew(inner, outer, spin, time), on HIP device `device`; NULL if it cannot run there (logged). */
ArtfDeviceFluid* artfDeviceFluidNew(float inner, float outer, float spin, double time, int device);
void artfDeviceFluidFree(ArtfDeviceFluid* f);
/* As artfFluidAdvance: true if the field changed. */
bool artfDeviceFluidAdvance(ArtfDeviceFluid* f, double time);
/* The field in device memory, ARTF_FLUID_WIDTH * ARTF_FLUID_HEIGHT floats, ready when New or Advance
 * returns. */
const float* artfDeviceFluidField(const ArtfDeviceFluid* f);
/* The field copied to the host. */
bool artfDeviceFluidRead(const ArtfDeviceFluid* f, float* field);
double artfDeviceFluidTime(const ArtfDeviceFluid* f);

/* Where a scene keeps its black hole's fluid: device >= 0 steps it on that HIP device as an
 * ArtfDeviceFluid (same bits, same calls: artfSceneSetBlackHole makes or advances it), which a HIP
 * renderer on the same device then reads in place, with no host step and no upload; -1 (the default)
 * keeps it on the host, as the reference does. A scene that already has a disc starts its fluid again
 * at the description's time. Falls back to the host, logged, if the device cannot hold it. */
typedef struct ArtfScene ArtfScene;
void artfSceneSetFluidOnDevice(ArtfScene* scene, int device);

#ifdef __cplusplus
}
#endif

#endif
