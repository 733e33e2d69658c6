"""ctypes layouts of the C structs in include/artful/*.h and include/artful_hip.h.

Only layouts live here: the calls are bound in artful_amd.native. ArtfBlackHoleDesc, ArtfCamera and
ArtfHoleFrame keep the byte layouts of Artful 0.4.0 (include/artful/scene.h:314-322, 470-490;
src/scene/sceneInternal.h:302-319), so one Python layout serves the reference's library too.
"""
from __future__ import annotations

import ctypes as C

F4 = C.c_float * 4
D4 = C.c_double * 4


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def __init__(self, x=0.0, y=0.0, z=0.0):
        super().__init__(x, y, z)

    def tolist(self):
        return [self.x, self.y, self.z]


class Camera(C.Structure):
    _fields_ = [("position", Vec3), ("target", Vec3), ("up", Vec3), ("fovY", C.c_float), ("near_", C.c_float)]


class BlackHoleDesc(C.Structure):
    _fields_ = [("position", Vec3), ("axis", Vec3), ("mass", C.c_float), ("spin", C.c_float),
                ("discInner", C.c_float), ("discOuter", C.c_float), ("discHeight", C.c_float),
                ("discThickness", C.c_float), ("discGain", C.c_float), ("doppler", C.c_float),
                ("stars", C.c_float), ("bloom", C.c_float), ("exposure", C.c_float), ("time", C.c_float),
                ("shutter", C.c_float), ("velocity", Vec3), ("samples", C.c_int)]


class HoleFrame(C.Structure):
    _fields_ = [(n, F4) for n in ("er", "etheta", "ephi", "right", "up", "forward", "metric", "place", "spin",
                                  "flow", "boost", "disc", "look", "image")] + \
               [("size", C.c_uint32 * 4), ("bloom", (C.c_uint32 * 4) * 8)]


class HoleFrame64(C.Structure):
    _fields_ = [(n, D4) for n in ("er", "etheta", "ephi", "right", "up", "forward", "metric", "place", "spin",
                                  "boost", "disc")]


class HoleTracing(C.Structure):
    _fields_ = [("precision", C.c_int), ("disc", C.c_int), ("emissivity", C.c_float), ("maxSteps", C.c_int),
                ("escapeRadius", C.c_float), ("jitter", C.c_uint32), ("bloomLevels", C.c_int), ("uncounted", C.c_int)]


class HoleStats(C.Structure):
    _fields_ = [("traceMs", C.c_double), ("resolveMs", C.c_double), ("bloomMs", C.c_double), ("toneMs", C.c_double),
                ("uploadMs", C.c_double), ("rays", C.c_uint64), ("steps", C.c_uint64), ("laneSlots", C.c_uint64),
                ("bloomLevels", C.c_int), ("frames", C.c_uint32), ("spanMs", C.c_double)]


class FrameStats(C.Structure):
    _fields_ = [("cpuMs", C.c_double), ("cullMs", C.c_double), ("geometryMs", C.c_double), ("rasterMs", C.c_double),
                ("lensMs", C.c_double), ("instancesDrawn", C.c_uint32), ("trianglesBinned", C.c_uint32),
                ("tilesSkipped", C.c_uint32), ("instancesHidden", C.c_uint32), ("traceSamples", C.c_uint32),
                ("gpuMs", C.c_double), ("shadowMapsDrawn", C.c_uint32)]


class Surface(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_void_p), ("b", C.c_void_p)]


class RendererOptions(C.Structure):
    _fields_ = [("backend", C.c_int), ("surface", Surface), ("width", C.c_int), ("height", C.c_int),
                ("noVsync", C.c_bool), ("threads", C.c_int), ("simd", C.c_int), ("antialias", C.c_bool),
                ("hdr", C.c_bool), ("compressTextures", C.c_bool), ("pipelineCache", C.c_char_p)]


class Overlay(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int),
                ("partial", C.c_bool), ("damage", C.c_void_p), ("damageCount", C.c_size_t)]


class HoleGradient(C.Structure):
    _fields_ = [("spin", C.c_double), ("position", C.c_double * 3), ("discInner", C.c_double), ("discOuter", C.c_double),
                ("emissivity", C.c_double), ("gain", C.c_double), ("frame", C.c_double * 26), ("spinDirect", C.c_double)]


class HipTraceArgs(C.Structure):
    _fields_ = [("frame", HoleFrame), ("frame64", HoleFrame64), ("map", C.c_void_p), ("fluid", C.c_void_p),
                ("out", C.c_void_p * 3), ("counters", C.c_void_p), ("precision", C.c_int), ("maxSteps", C.c_int),
                ("escapeRadius", C.c_float), ("disc", C.c_int), ("emissivity", C.c_float), ("jitter", C.c_uint32),
                ("rowBegin", C.c_int), ("rowEnd", C.c_int), ("startStamp", C.c_void_p), ("spanStamp", C.c_void_p)]


class HipFrameArgs(C.Structure):
    _fields_ = [("trace", HipTraceArgs), ("hdr", C.c_void_p * 3), ("bloom", C.c_void_p * 3), ("bgra", C.c_void_p), ("report", C.c_void_p),
                ("postStream", C.c_void_p), ("handoff", C.c_void_p), ("done", C.c_void_p)]


REPORT_WORDS = 7  # ARTF_HIP_REPORT_WORDS


class HipAdjointArgs(C.Structure):
    _fields_ = [("trace", HipTraceArgs), ("seed", C.c_void_p * 3), ("partials", C.c_void_p)]


class HipView(C.Structure):
    _fields_ = [("frame", HoleFrame), ("frame64", HoleFrame64), ("hdr", C.c_void_p * 3), ("rays", C.c_void_p * 3), ("maxSteps", C.c_int32),
                ("escapeRadius", C.c_float), ("disc", C.c_int32), ("emissivity", C.c_float), ("jitter", C.c_uint32), ("tilesX", C.c_uint32),
                ("tilesY", C.c_uint32), ("rayCount", C.c_uint32)]


class HipSchedule(C.Structure):
    _fields_ = [("workgroups", C.c_int32), ("chunkTiles", C.c_int32), ("refill", C.c_int32), ("stepsPerCheck", C.c_int32)]


class HipBatchArgs(C.Structure):
    _fields_ = [("views", C.c_void_p), ("firstTile", C.c_void_p), ("viewCount", C.c_int32), ("tileBegin", C.c_uint32),
                ("tileEnd", C.c_uint32),
                ("precision", C.c_int32), ("map", C.c_void_p), ("rings", C.c_void_p), ("fluids", C.c_void_p),
                ("queue", C.c_void_p), ("counters", C.c_void_p), ("stamps", C.c_void_p), ("schedule", HipSchedule),
                ("dragInRange", C.c_int32)]


class HoleView(C.Structure):
    _fields_ = [("hole", BlackHoleDesc), ("camera", Camera), ("width", C.c_int32), ("height", C.c_int32)]


class HoleSchedule(C.Structure):
    _fields_ = [("workgroups", C.c_int32), ("chunkTiles", C.c_int32), ("refill", C.c_int32), ("stepsPerCheck", C.c_int32),
                ("launchTiles", C.c_int32)]


class HoleBatchStats(C.Structure):
    _fields_ = [("batchMs", C.c_double), ("traceMs", C.c_double), ("rays", C.c_uint64), ("steps", C.c_uint64),
                ("laneSlots", C.c_uint64), ("views", C.c_int32), ("schedule", HoleSchedule)]


# enum values (include/artful/renderer.h, include/artful/hole.h)
BACKEND_AUTO, BACKEND_VULKAN, BACKEND_SOFTWARE, BACKEND_METAL, BACKEND_HIP = range(5)
HOLE_WORKGROUP_A_TILE = -1  # ARTF_HOLE_WORKGROUP_A_TILE: the batch with a workgroup a tile, no queue
PRECISION = {"fp32": 0, "fp32_fast": 1, "mixed": 2, "fp64": 3}
DISC = {"painted": 0, "analytic": 1}

# sizes from include/artful/hole.h
MAP_WIDTH, MAP_HEIGHT = 2048, 1024
DUST_WIDTH, DUST_HEIGHT = 1024, 256
FLUID_WIDTH, FLUID_HEIGHT = 1024, 256
COLORS = 256
SKY_WIDTH, SKY_HEIGHT = 1024, 512
ROWS_AT = MAP_WIDTH * MAP_HEIGHT
COLORS_AT = ROWS_AT + MAP_HEIGHT * 2
SKY_AT = COLORS_AT + COLORS * 3
GLOW_AT = SKY_AT + SKY_WIDTH * SKY_HEIGHT
DUST_AT = GLOW_AT + MAP_HEIGHT
MAP_FLOATS = DUST_AT + DUST_WIDTH * DUST_HEIGHT

assert C.sizeof(BlackHoleDesc) == 92 and C.sizeof(Camera) == 44 and C.sizeof(HoleFrame) == 368
assert C.sizeof(HipView) == 800 and C.sizeof(HoleView) == 144
