"""ctypes binding of the oracle (oracle/liboracle.so) and, where it was built, of the reference
library (oracle/_ref/libartful_ref.so). Test infrastructure only: imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, to test the HIP path against."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from artful_amd import abi

HERE = Path(__file__).resolve().parent
ORACLE = HERE / "liboracle.so"
REFERENCE = HERE / "_ref" / "libartful_ref.so"


class Law(C.Structure):
    _fields_ = [("maxSteps", C.c_int), ("escapeRadius", C.c_float), ("disc", C.c_int), ("emissivity", C.c_float),
                ("jitter", C.c_uint32)]


_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE.exists():
            raise RuntimeError(f"{ORACLE} is not built: run python oracle/build.py")
        L = C.CDLL(str(ORACLE))
        P = C.c_void_p
        L.oracleTrace.argtypes = [C.POINTER(abi.HoleFrame), C.POINTER(abi.HoleFrame64), C.POINTER(Law), C.c_int, P, P, P, P, P, P,
                                  C.c_int, C.c_int, C.c_int]
        L.oracleTrace.restype = None
        L.oracleRay.argtypes = [C.POINTER(abi.HoleFrame), C.POINTER(abi.HoleFrame64), C.POINTER(Law), C.c_int, P, P, C.c_double,
                                C.c_double, C.c_float, P, C.POINTER(C.c_int), P]
        L.oracleRay.restype = None
        L.oracleBloom.argtypes = [C.POINTER(abi.HoleFrame), P * 3, P * 3]
        L.oracleBloom.restype = None
        L.oracleToneMap.argtypes = [C.POINTER(abi.HoleFrame), P * 3, P * 3, P]
        L.oracleToneMap.restype = None
        _oracle = L
    return _oracle


def law(max_steps=0, escape_radius=0.0, disc="painted", emissivity=3.0, jitter=0) -> Law:
    return Law(max_steps, escape_radius, abi.DISC[disc] if isinstance(disc, str) else disc, emissivity, jitter)


def trace(frame: abi.HoleFrame, map_: np.ndarray, fluid: np.ndarray, precision: int = 32, frame64: abi.HoleFrame64 | None = None,
          lw: Law | None = None, rows: tuple[int, int] | None = None, threads: int = 0, steps: bool = False):
    """HDR (3, H, W) float32, and with steps=True the per-ray step counts (side^2, H, W)."""
    W, H, S = frame.size[0], frame.size[1], frame.size[2]
    lw = lw or law()
    map_ = np.ascontiguousarray(map_, np.float32)
    fluid = np.ascontiguousarray(fluid, np.float32)
    out = np.zeros((3, H, W), np.float32)
    st = np.zeros((S * S, H, W), np.uint32) if steps else None
    r0, r1 = rows if rows else (0, H)
    oracle().oracleTrace(C.byref(frame), C.byref(frame64) if frame64 is not None else None, C.byref(lw), precision,
                         map_.ctypes.data, fluid.ctypes.data, out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                         st.ctypes.data if st is not None else None, r0, r1, threads)
    return (out, st) if steps else out


def ray(frame, map_, fluid, nx, ny, time, precision=32, frame64=None, lw=None):
    """(rgb, steps, path): path = where the ray left on the unit sphere (3), its drag angle, 1.0 if it
    left, and the transmittance."""
    rgb = np.zeros(3, np.float32)
    path = np.zeros(6, np.float64)
    n = C.c_int()
    oracle().oracleRay(C.byref(frame), C.byref(frame64) if frame64 is not None else None, C.byref(lw or law()), precision,
                       np.ascontiguousarray(map_, np.float32).ctypes.data, np.ascontiguousarray(fluid, np.float32).ctypes.data,
                       nx, ny, time, rgb.ctypes.data, C.byref(n), path.ctypes.data)
    return rgb, n.value, path


def bloom(frame: abi.HoleFrame, hdr: np.ndarray, bloom_pixels: int) -> np.ndarray:
    hdr = np.ascontiguousarray(hdr, np.float32)
    out = np.zeros((3, bloom_pixels), np.float32)
    P3 = C.c_void_p * 3
    oracle().oracleBloom(C.byref(frame), P3(*[hdr[c].ctypes.data for c in range(3)]), P3(*[out[c].ctypes.data for c in range(3)]))
    return out


def tone_map(frame: abi.HoleFrame, hdr: np.ndarray, bl: np.ndarray) -> np.ndarray:
    hdr = np.ascontiguousarray(hdr, np.float32)
    bl = np.ascontiguousarray(bl, np.float32)
    out = np.zeros((frame.size[1], frame.size[0]), np.uint32)
    P3 = C.c_void_p * 3
    oracle().oracleToneMap(C.byref(frame), P3(*[hdr[c].ctypes.data for c in range(3)]), P3(*[bl[c].ctypes.data for c in range(3)]),
                           out.ctypes.data)
    return out


def render(frame, map_, fluid, bloom_pixels, precision=32, frame64=None, lw=None):
    """The whole CPU path: HDR, bloom, BGRA."""
    hdr = trace(frame, map_, fluid, precision, frame64, lw)
    bl = bloom(frame, hdr, bloom_pixels)
    return hdr, bl, tone_map(frame, hdr, bl)


# ---- the reference's own library (built in the container that has the reference) ----

def reference_available() -> bool:
    return REFERENCE.exists()


def reference() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REFERENCE.exists():
            raise RuntimeError(f"{REFERENCE} is not built (it needs the reference tree: python oracle/build.py)")
        L = C.CDLL(str(REFERENCE))
        P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
        sig = {
            "refSceneNew": ([C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera)], P),
            "refSceneSet": ([P, C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera)], None),
            "refSceneFree": ([P], None), "refSceneDesc": ([P, C.POINTER(abi.BlackHoleDesc)], None),
            "refMap": ([P], C.POINTER(C.c_float)), "refFluid": ([P], C.POINTER(C.c_float)), "refFluidTime": ([P], D),
            "refMapVersion": ([P], C.c_ulonglong), "refFluidVersion": ([P], C.c_ulonglong),
            "refMapFloats": ([], I), "refHoleFrameBytes": ([], I),
            "refFrame": ([P, I, I, C.POINTER(abi.HoleFrame)], C.c_uint32),
            "refIsco": ([F], F), "refOrbitSpeed": ([F, F], F), "refDefaults": ([C.POINTER(abi.BlackHoleDesc)], None),
            "refFluidNew": ([F, F, F, D], P), "refFluidAdvance": ([P, D], I), "refFluidField": ([P], C.POINTER(C.c_float)),
            "refFluidFieldTime": ([P], D), "refFluidFree": ([P], None),
            "refLens": ([P, I, I, I, P, P, P, P], None),
            "refDraw": ([P, I, I, I, I, P], D),
            "refRendererNew": ([I, I, I, I], P), "refRendererDraw": ([P, P], I), "refRendererRead": ([P, P], None),
            "refRendererStats": ([P, C.POINTER(D), C.POINTER(D), C.POINTER(D)], None), "refRendererFree": ([P], None),
            "refRendererDescription": ([P], C.c_char_p),
            "refLayoutCount": ([], I), "refLayoutType": ([I], C.c_char_p), "refLayoutField": ([I], C.c_char_p),
            "refLayoutAt": ([I], C.c_long), "refEnumValue": ([I], I),
        }
        for name, (args, ret) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ret
        _ref = L
    return _ref


class RefScene:
    """A scene in the reference library."""

    def __init__(self, desc: abi.BlackHoleDesc, cam: abi.Camera):
        self.L = reference()
        self.ptr = self.L.refSceneNew(C.byref(desc), C.byref(cam))

    def set(self, desc: abi.BlackHoleDesc, cam: abi.Camera | None = None) -> None:
        self.L.refSceneSet(self.ptr, C.byref(desc), C.byref(cam) if cam is not None else None)

    def map(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.L.refMap(self.ptr), shape=(abi.MAP_FLOATS,)).copy()

    def fluid(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.L.refFluid(self.ptr), shape=(abi.FLUID_WIDTH * abi.FLUID_HEIGHT,)).copy()

    def frame(self, width: int, height: int) -> tuple[abi.HoleFrame, int]:
        f = abi.HoleFrame()
        n = self.L.refFrame(self.ptr, width, height, C.byref(f))
        return f, n

    def lens(self, width: int, height: int, isa: int = 2) -> np.ndarray:
        out = np.zeros((3, height, width), np.float32)
        self.L.refLens(self.ptr, width, height, isa, None, out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data)
        return out

    def draw(self, width: int, height: int, simd: int = 0, threads: int = 1) -> np.ndarray:
        out = np.zeros((height, width), np.uint32)
        self.L.refDraw(self.ptr, width, height, simd, threads, out.ctypes.data)
        return out

    def close(self) -> None:
        if self.ptr:
            self.L.refSceneFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
