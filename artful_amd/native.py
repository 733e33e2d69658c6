"""Python binding of libartful_hip.so (the C library: scene, fluid, HIP renderer, kernel C ABI).

A thin ctypes layer for tests, the bench and scripts; the library itself is C and HIP. Loading fails
loudly when the library has not been built (python -m artful_amd.build): there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import abi

LIB_PATH = Path(os.environ.get("ARTF_HIP_LIBRARY") or Path(__file__).resolve().parent / "libartful_hip.so")
_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run python -m artful_amd.build")
    L = C.CDLL(str(LIB_PATH))
    P, I, F, D, U32, U64, B = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_uint32, C.c_uint64, C.c_bool
    sig = {
        "artfSceneNew": ([], P), "artfSceneFree": ([P], None),
        "artfSceneSetCamera": ([P, C.POINTER(abi.Camera)], None),
        "artfSceneSetBlackHole": ([P, C.POINTER(abi.BlackHoleDesc)], None),
        "artfBlackHoleDefaults": ([], abi.BlackHoleDesc), "artfBlackHoleOrbitSpeed": ([F, F], F),
        "artfHoleIsco": ([F], F),
        "artfSceneHoleDesc": ([P, C.POINTER(abi.BlackHoleDesc)], B),
        "artfSceneHoleMap": ([P, C.POINTER(U64)], C.POINTER(C.c_float)),
        "artfSceneHoleFluid": ([P, C.POINTER(U64), C.POINTER(D)], C.POINTER(C.c_float)),
        "artfSceneHoleFrame": ([P, I, I, I, C.POINTER(abi.HoleFrame)], U32),
        "artfSceneHoleFrame64": ([P, I, I, C.POINTER(abi.HoleFrame64)], None),
        "artfHoleFrameFor": ([C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera), I, I, C.POINTER(abi.HoleFrame),
                              C.POINTER(abi.HoleFrame64)], None),
        "artfFluidNew": ([F, F, F, D], P), "artfFluidFree": ([P], None), "artfFluidAdvance": ([P, D], B),
        "artfFluidField": ([P], C.POINTER(C.c_float)), "artfFluidTime": ([P], D),
        "artfRendererDefaults": ([I, I], abi.RendererOptions),
        "artfRendererNew": ([C.POINTER(abi.RendererOptions)], P), "artfRendererFree": ([P], None),
        "artfHipRendererNew": ([C.POINTER(abi.RendererOptions), I], P),
        "artfSceneSetBackground": ([P, abi.Vec3], None), "artfSceneSetBackgroundAlpha": ([P, F], None),
        "artfRendererBackend": ([P], I), "artfRendererDescription": ([P], C.c_char_p),
        "artfRendererResize": ([P, I, I], None), "artfRendererSetOverlay": ([P, C.POINTER(abi.Overlay)], None),
        "artfRendererDraw": ([P, P], B), "artfRendererReadPixels": ([P, P], B), "artfRendererPixels": ([P], P),
        "artfRendererStats": ([P, C.POINTER(abi.FrameStats)], None),
        "artfRendererReadHdr": ([P, P, P, P], B),
        "artfHoleTracingDefaults": ([], abi.HoleTracing),
        "artfRendererSetHoleTracing": ([P, C.POINTER(abi.HoleTracing)], None),
        "artfRendererHoleStats": ([P, C.POINTER(abi.HoleStats)], B),
        "artfBackendName": ([I], C.c_char_p),
        "artfHipDeviceCount": ([], I), "artfHipSetDevice": ([I], I),
        "artfHipDeviceInfo": ([I, C.c_char_p, C.c_size_t, C.POINTER(I)], I),
        "artfHipMalloc": ([C.c_size_t], P), "artfHipFree": ([P], None),
        "artfHipUpload": ([P, P, C.c_size_t, P], I), "artfHipDownload": ([P, P, C.c_size_t, P], I),
        "artfHipMemset": ([P, I, C.c_size_t, P], I), "artfHipCopy": ([P, P, C.c_size_t, P], I),
        "artfHipStreamCreate": ([], P), "artfHipStreamDestroy": ([P], None), "artfHipStreamSync": ([P], I), "artfHipDeviceSync": ([], I),
        "artfHipEventCreate": ([], P), "artfHipEventDestroy": ([P], None), "artfHipEventRecord": ([P, P], I),
        "artfHipEventMs": ([P, P], F), "artfHipErrorText": ([I], C.c_char_p),
        "artfHipHostMalloc": ([C.c_size_t], P), "artfHipHostFree": ([P], None),
        "artfHipDownloadAsync": ([P, P, C.c_size_t, P], I),
        "artfHipDrawFrame": ([C.POINTER(abi.HipFrameArgs), P], I),
        "artfHipDrawFrameStatic": ([C.POINTER(abi.HipFrameArgs), P], I),
        "artfHipEventCreateOrder": ([], P), "artfHipStreamWait": ([P, P], I), "artfHipEventDone": ([P], I),
        "artfHipTrace": ([C.POINTER(abi.HipTraceArgs), P], I),
        "artfHipResolve": ([C.POINTER(abi.HoleFrame), C.c_void_p * 3, C.c_void_p * 3, P], I),
        "artfHipBloom": ([C.POINTER(abi.HoleFrame), C.c_void_p * 3, C.c_void_p * 3, P], I),
        "artfHipToneMap": ([C.POINTER(abi.HoleFrame), C.c_void_p * 3, C.c_void_p * 3, P, P], I),
        "artfHipAdjointBlocks": ([C.POINTER(abi.HoleFrame)], I),
        "artfHipAdjoint": ([C.POINTER(abi.HipAdjointArgs), P], I),
        "artfHoleAdjointNew": ([I, I, I], P), "artfHoleAdjointFree": ([P], None),
        "artfHoleAdjointRender": ([P, C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera), C.POINTER(abi.HoleTracing), P], B),
        "artfHoleAdjointGradient": ([P, C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera), C.POINTER(abi.HoleTracing), P, P,
                                     C.POINTER(abi.HoleGradient)], B),
        "artfHoleAdjointTimes": ([P, C.POINTER(D), C.POINTER(D)], None),
        "artfHipBatchReset": ([C.POINTER(abi.HipBatchArgs), P], I),
        "artfHipTraceViews": ([C.POINTER(abi.HipBatchArgs), P], I),
        "artfHipViewsOccupancy": ([I], I), "artfHipWallClockHz": ([I], D),
        "artfHipGraphBegin": ([P], I), "artfHipGraphEnd": ([P, C.POINTER(P)], I),
        "artfHipGraphLaunch": ([P, P], I), "artfHipGraphFree": ([P], None),
        "artfHoleBatchNew": ([C.POINTER(abi.HoleView), I, C.POINTER(abi.HoleTracing), I], P),
        "artfHoleBatchFree": ([P], None), "artfHoleBatchDraw": ([P], B),
        "artfHoleBatchReadHdr": ([P, I, P], B),
        "artfHoleBatchStats": ([P, C.POINTER(abi.HoleBatchStats)], B),
        "artfHoleBatchLatency": ([P, P], B),
        "artfHoleBatchSetSchedule": ([P, C.POINTER(abi.HoleSchedule)], B),
        "artfHoleBatchTune": ([P, I, C.POINTER(abi.HoleSchedule)], B),
        "artfDeviceFluidNew": ([F, F, F, D, I], P), "artfDeviceFluidFree": ([P], None),
        "artfDeviceFluidAdvance": ([P, D], B), "artfDeviceFluidField": ([P], P),
        "artfDeviceFluidRead": ([P, P], B), "artfDeviceFluidTime": ([P], D),
        "artfSceneSetDiscOnDevice": ([P, I], None),
        "artfDeviceMapNew": ([I], P), "artfDeviceMapFree": ([P], None), "artfDeviceMapRings": ([P, F, F, F, I], B),
    }
    for name, (args, ret) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ret
    L.artfHipCallsLeft = C.c_int.in_dll(L, "artfHipCallsLeft")
    _lib = L
    return L


def declared_symbols() -> list[str]:
    """The functions include/*.h declares (for the export check)."""
    import re
    root = Path(__file__).resolve().parent.parent / "include"
    names = set()
    for h in list(root.glob("*.h")) + list(root.glob("artful/*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(artf\w+)\s*\(", text, flags=re.M):
            if "ARTF_INLINE" in m.group(0) or "typedef" in m.group(0) or "#define" in m.group(0):
                continue
            names.add(m.group(1))
    return sorted(names)


def defaults() -> abi.BlackHoleDesc:
    return lib().artfBlackHoleDefaults()


def camera(position, target=(0, 0, 0), up=(0, 1, 0), fov_y=1.0, near=0.1) -> abi.Camera:
    return abi.Camera(abi.Vec3(*position), abi.Vec3(*target), abi.Vec3(*up), fov_y, near)


def hole_desc(**kw) -> abi.BlackHoleDesc:
    d = defaults()
    for k, v in kw.items():
        if k in ("position", "axis", "velocity"):
            setattr(d, k, abi.Vec3(*v))
        else:
            setattr(d, k, v)
    return d


class Scene:
    """A scene with a camera and a black hole (artfSceneNew ... artfSceneFree)."""

    def __init__(self, desc: abi.BlackHoleDesc | None = None, cam: abi.Camera | None = None):
        self._L = lib()
        self.ptr = self._L.artfSceneNew()
        if cam is not None:
            self.set_camera(cam)
        if desc is not None:
            self.set_hole(desc)

    def set_camera(self, cam: abi.Camera) -> None:
        self._L.artfSceneSetCamera(self.ptr, C.byref(cam))

    def set_background(self, color, alpha: float | None = None) -> None:
        """artfSceneSetBackground (0..1 RGB), and artfSceneSetBackgroundAlpha when alpha is given."""
        self._L.artfSceneSetBackground(self.ptr, abi.Vec3(*color))
        if alpha is not None:
            self._L.artfSceneSetBackgroundAlpha(self.ptr, alpha)

    def set_disc_on_device(self, device: int) -> None:
        """Build the disc's map and step its fluid on HIP device `device` (-1: on the host, the default)."""
        self._L.artfSceneSetDiscOnDevice(self.ptr, device)

    def set_hole(self, desc: abi.BlackHoleDesc | None) -> None:
        self._L.artfSceneSetBlackHole(self.ptr, C.byref(desc) if desc is not None else None)

    def desc(self) -> abi.BlackHoleDesc | None:
        d = abi.BlackHoleDesc()
        return d if self._L.artfSceneHoleDesc(self.ptr, C.byref(d)) else None

    def map(self) -> np.ndarray | None:
        v = C.c_uint64()
        p = self._L.artfSceneHoleMap(self.ptr, C.byref(v))
        if not p:  # no disc, or its map on a device
            return None
        return np.ctypeslib.as_array(p, shape=(abi.MAP_FLOATS,)).copy()

    def map_version(self) -> int:
        v = C.c_uint64()
        self._L.artfSceneHoleMap(self.ptr, C.byref(v))
        return v.value

    def fluid(self) -> tuple[np.ndarray, int, float]:
        v, t = C.c_uint64(), C.c_double()
        p = self._L.artfSceneHoleFluid(self.ptr, C.byref(v), C.byref(t))
        if not p:  # no disc, or its fluid on a device
            return None, v.value, t.value
        return np.ctypeslib.as_array(p, shape=(abi.FLUID_WIDTH * abi.FLUID_HEIGHT,)).copy(), v.value, t.value

    def frame(self, width: int, height: int, levels: int = 0) -> tuple[abi.HoleFrame, int]:
        f = abi.HoleFrame()
        n = self._L.artfSceneHoleFrame(self.ptr, width, height, levels, C.byref(f))
        return f, n

    def frame64(self, width: int, height: int) -> abi.HoleFrame64:
        f = abi.HoleFrame64()
        self._L.artfSceneHoleFrame64(self.ptr, width, height, C.byref(f))
        return f

    def close(self) -> None:
        if self.ptr:
            self._L.artfSceneFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def frames_for(desc, cam, width, height):
    """(ArtfHoleFrame, ArtfHoleFrame64) for a description and camera, without a scene."""
    f, f64 = abi.HoleFrame(), abi.HoleFrame64()
    lib().artfHoleFrameFor(C.byref(desc), C.byref(cam), width, height, C.byref(f), C.byref(f64))
    return f, f64


def black_body_colors() -> np.ndarray:
    """The map's colour table (what the analytic disc reads), from a scene built once."""
    global _colors
    if _colors is None:
        s = Scene(hole_desc(), camera((0, 0, 30)))
        _colors = s.map()[abi.COLORS_AT:abi.SKY_AT].copy()
        s.close()
    return _colors


_colors = None


def device_count() -> int:
    return lib().artfHipDeviceCount()


class Renderer:
    """artfRendererNew with the HIP backend; raises when no GPU can run it."""

    def __init__(self, width: int, height: int, device: int = 0):
        self._L = lib()
        o = self._L.artfRendererDefaults(width, height)
        o.backend = abi.BACKEND_HIP
        self.ptr = self._L.artfHipRendererNew(C.byref(o), device)
        if not self.ptr:
            raise RuntimeError(f"artfHipRendererNew: HIP device {device} cannot run the renderer here")
        self.width, self.height = width, height

    def description(self) -> str:
        return self._L.artfRendererDescription(self.ptr).decode()

    def set_tracing(self, precision="fp32", disc="painted", emissivity=3.0, max_steps=0, escape_radius=0.0,
                    jitter=0, bloom_levels=0, count_steps=True) -> None:
        t = self._L.artfHoleTracingDefaults()
        t.precision = abi.PRECISION[precision] if isinstance(precision, str) else precision
        t.disc = abi.DISC[disc] if isinstance(disc, str) else disc
        t.emissivity, t.maxSteps, t.escapeRadius = emissivity, max_steps, escape_radius
        t.jitter, t.bloomLevels, t.uncounted = jitter, bloom_levels, 0 if count_steps else 1
        self._L.artfRendererSetHoleTracing(self.ptr, C.byref(t))

    def resize(self, width: int, height: int) -> None:
        self._L.artfRendererResize(self.ptr, width, height)
        self.width, self.height = width, height

    def draw(self, scene: Scene) -> bool:
        return bool(self._L.artfRendererDraw(self.ptr, scene.ptr))

    def read_pixels(self) -> np.ndarray:
        out = np.zeros((self.height, self.width), np.uint32)
        if not self._L.artfRendererReadPixels(self.ptr, out.ctypes.data):
            raise RuntimeError("artfRendererReadPixels failed")
        return out

    def read_hdr(self) -> np.ndarray:
        out = np.zeros((3, self.height, self.width), np.float32)
        if not self._L.artfRendererReadHdr(self.ptr, out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data):
            raise RuntimeError("artfRendererReadHdr failed")
        return out

    def stats(self) -> abi.FrameStats:
        s = abi.FrameStats()
        self._L.artfRendererStats(self.ptr, C.byref(s))
        return s

    def hole_stats(self) -> abi.HoleStats:
        s = abi.HoleStats()
        self._L.artfRendererHoleStats(self.ptr, C.byref(s))
        return s

    def set_overlay(self, pixels: np.ndarray | None) -> None:
        if pixels is None:
            self._L.artfRendererSetOverlay(self.ptr, None)
            return
        self._overlay = np.ascontiguousarray(pixels, np.uint32)
        o = abi.Overlay(self._overlay.ctypes.data, self._overlay.shape[1], self._overlay.shape[0], 0, False, None, 0)
        self._L.artfRendererSetOverlay(self.ptr, C.byref(o))

    def close(self) -> None:
        if self.ptr:
            self._L.artfRendererFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tracing(precision="fp32", disc="analytic", emissivity=3.0, max_steps=0, escape_radius=0.0, jitter=0, bloom_levels=0, count_steps=True):
    t = lib().artfHoleTracingDefaults()
    t.precision = abi.PRECISION[precision] if isinstance(precision, str) else precision
    t.disc = abi.DISC[disc] if isinstance(disc, str) else disc
    t.emissivity, t.maxSteps, t.escapeRadius, t.jitter, t.bloomLevels = emissivity, max_steps, escape_radius, jitter, bloom_levels
    t.uncounted = 0 if count_steps else 1
    return t


class HoleAdjoint:
    """The differentiable analytic-disc renderer (include/artful/adjoint.h)."""

    def __init__(self, width: int, height: int, device: int = 0):
        self._L = lib()
        self.ptr = self._L.artfHoleAdjointNew(width, height, device)
        if not self.ptr:
            raise RuntimeError("artfHoleAdjointNew: no HIP device")
        self.width, self.height = width, height

    def render(self, desc, cam, trc) -> np.ndarray:
        out = np.zeros((3, self.height, self.width), np.float32)
        if not self._L.artfHoleAdjointRender(self.ptr, C.byref(desc), C.byref(cam), C.byref(trc), out.ctypes.data):
            raise RuntimeError("artfHoleAdjointRender failed")
        return out

    def gradient(self, desc, cam, trc, seed: np.ndarray):
        seed = np.ascontiguousarray(seed, np.float32)
        assert seed.shape == (3, self.height, self.width)
        hdr = np.zeros((3, self.height, self.width), np.float32)
        g = abi.HoleGradient()
        if not self._L.artfHoleAdjointGradient(self.ptr, C.byref(desc), C.byref(cam), C.byref(trc), seed.ctypes.data, hdr.ctypes.data, C.byref(g)):
            raise RuntimeError("artfHoleAdjointGradient failed")
        return g, hdr

    def times(self) -> tuple[float, float]:
        f, a = C.c_double(), C.c_double()
        self._L.artfHoleAdjointTimes(self.ptr, C.byref(f), C.byref(a))
        return f.value, a.value

    def close(self) -> None:
        if self.ptr:
            self._L.artfHoleAdjointFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calls_left(n: int) -> None:
    """Test hook: device calls that succeed before the device is reported lost (-1: no limit)."""
    lib().artfHipCallsLeft.value = n


class HoleBatch:
    """Many views traced by one launch (include/artful/batch.h). views: (desc, camera, width,
    height) tuples; tracing: a HoleTracing (artful_amd.native.tracing) for all of them."""

    def __init__(self, views, trc=None, device: int = 0):
        self._L = lib()
        arr = (abi.HoleView * len(views))()
        self.sizes = []
        for i, (desc, cam, w, h) in enumerate(views):
            arr[i].hole, arr[i].camera, arr[i].width, arr[i].height = desc, cam, w, h
            self.sizes.append((w, h))
        self.ptr = self._L.artfHoleBatchNew(arr, len(views), C.byref(trc) if trc is not None else None, device)
        if not self.ptr:
            raise RuntimeError("artfHoleBatchNew failed")

    def draw(self) -> None:
        if not self._L.artfHoleBatchDraw(self.ptr):
            raise RuntimeError("artfHoleBatchDraw failed")

    def read_hdr(self, view: int) -> np.ndarray:
        w, h = self.sizes[view]
        out = np.zeros((3, h, w), np.float32)
        if not self._L.artfHoleBatchReadHdr(self.ptr, view, out.ctypes.data):
            raise RuntimeError("artfHoleBatchReadHdr failed")
        return out

    def stats(self) -> abi.HoleBatchStats:
        s = abi.HoleBatchStats()
        if not self._L.artfHoleBatchStats(self.ptr, C.byref(s)):
            raise RuntimeError("artfHoleBatchStats failed")
        return s

    def latency(self) -> np.ndarray:
        out = np.zeros(len(self.sizes), np.float64)
        if not self._L.artfHoleBatchLatency(self.ptr, out.ctypes.data):
            raise RuntimeError("artfHoleBatchLatency failed")
        return out

    def set_schedule(self, workgroups=0, chunk_tiles=0, refill=0, steps_per_check=0, launch_tiles=0) -> None:
        s = abi.HoleSchedule(workgroups, chunk_tiles, refill, steps_per_check, launch_tiles)
        if not self._L.artfHoleBatchSetSchedule(self.ptr, C.byref(s)):
            raise RuntimeError("artfHoleBatchSetSchedule failed")

    def tune(self, rounds: int = 0) -> abi.HoleSchedule:
        s = abi.HoleSchedule()
        if not self._L.artfHoleBatchTune(self.ptr, rounds, C.byref(s)):
            raise RuntimeError("artfHoleBatchTune failed")
        return s

    def close(self) -> None:
        if self.ptr:
            self._L.artfHoleBatchFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceFluid:
    """The disc fluid stepped on the GPU (include/artful/device_disc.h)."""

    def __init__(self, inner: float, outer: float, spin: float, time: float = 0.0, device: int = 0):
        self._L = lib()
        self.ptr = self._L.artfDeviceFluidNew(inner, outer, spin, time, device)
        if not self.ptr:
            raise RuntimeError("artfDeviceFluidNew failed")

    def advance(self, time: float) -> bool:
        return bool(self._L.artfDeviceFluidAdvance(self.ptr, time))

    def field(self) -> np.ndarray:
        out = np.zeros(abi.FLUID_WIDTH * abi.FLUID_HEIGHT, np.float32)
        if not self._L.artfDeviceFluidRead(self.ptr, out.ctypes.data):
            raise RuntimeError("artfDeviceFluidRead failed")
        return out

    def time(self) -> float:
        return self._L.artfDeviceFluidTime(self.ptr)

    def close(self) -> None:
        if self.ptr:
            self._L.artfDeviceFluidFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_map(spin: float, inner: float, outer: float, device: int = 0) -> np.ndarray:
    """A disc map painted on the GPU (artfDeviceMapNew, then this disc's rings), copied to the host."""
    L = lib()
    m = L.artfDeviceMapNew(device)
    if not m:
        raise RuntimeError("artfDeviceMapNew failed")
    try:
        if not L.artfDeviceMapRings(m, inner, outer, spin, device):
            raise RuntimeError("artfDeviceMapRings failed")
        out = np.zeros(abi.MAP_FLOATS, np.float32)
        if L.artfHipDownload(out.ctypes.data, m, out.nbytes, None):
            raise RuntimeError("download failed")
        return out
    finally:
        L.artfDeviceMapFree(m)

