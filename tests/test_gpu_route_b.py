"""Route B on the GPU: the HIP backend built into the reference's own tree (oracle/_ref/libartful_routeb.so:
the reference's scene code, backend dispatch and software renderer, compiled from its sources, with this
repository's hip_renderer.c and route-B scene port compiled against the reference's headers, and the HIP
kernels). The same reference scene is drawn through the reference's artfRendererDraw by its software
renderer and by artfHipNew's renderer: their frames must be equal bit for bit, with an overlay over them
(premultiplied pixels of every kind, alpha 0 with colour among them), without a hole (the scene's
background, at full and half alpha), and the HIP renderer's HDR equal to the reference's own lens kernel."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from artful_amd import abi
from tests.golden import cases

pytestmark = pytest.mark.gpu

ROUTE_B = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libartful_routeb.so"


@pytest.fixture(scope="module")
def ref(gpu):
    if not ROUTE_B.exists():
        pytest.fail(f"{ROUTE_B} is not built (make -C oracle/ref routeb, where the reference tree is)")
    L = C.CDLL(str(ROUTE_B))
    P, I, B = C.c_void_p, C.c_int, C.c_bool
    sig = {
        "refSceneNew": ([C.POINTER(abi.BlackHoleDesc), C.POINTER(abi.Camera)], P), "refSceneFree": ([P], None),
        "refLens": ([P, I, I, I, P, P, P, P], None),
        "artfRendererDefaults": ([I, I], abi.RendererOptions), "artfRendererNew": ([C.POINTER(abi.RendererOptions)], P),
        "artfHipNew": ([C.POINTER(abi.RendererOptions)], P), "artfRendererFree": ([P], None), "artfRendererBackend": ([P], I),
        "artfRendererDraw": ([P, P], B), "artfRendererReadPixels": ([P, P], B), "artfRendererResize": ([P, I, I], None),
        "artfRendererSetOverlay": ([P, C.POINTER(abi.Overlay)], None), "artfRendererReadHdr": ([P, P, P, P], B),
        "artfSceneSetBlackHole": ([P, C.POINTER(abi.BlackHoleDesc)], None), "artfSceneSetBackground": ([P, abi.Vec3], None),
        "artfSceneSetBackgroundAlpha": ([P, C.c_float], None), "artfRendererDescription": ([P], C.c_char_p),
    }
    for name, (args, ret) in sig.items():
        fn = getattr(L, name)
        fn.argtypes, fn.restype = args, ret
    return L


@pytest.fixture(scope="module")
def pair(ref):
    """A software renderer (the reference's own) and a HIP one (this backend, built into the reference)."""
    o = ref.artfRendererDefaults(64, 64)
    o.backend = abi.BACKEND_SOFTWARE
    o.threads = 4
    sw = ref.artfRendererNew(C.byref(o))
    hip = ref.artfHipNew(C.byref(o))
    assert sw and hip, "a renderer of each"
    assert ref.artfRendererBackend(hip) == abi.BACKEND_HIP and ref.artfRendererDescription(hip).decode().startswith("HIP")
    yield sw, hip
    ref.artfRendererFree(sw)
    ref.artfRendererFree(hip)


def frames(ref, pair, scene, w, h):
    out = []
    for r in pair:
        ref.artfRendererResize(r, w, h)
        assert ref.artfRendererDraw(r, scene)
        px = np.zeros((h, w), np.uint32)
        assert ref.artfRendererReadPixels(r, px.ctypes.data)
        out.append(px)
    return out


@pytest.mark.parametrize("name,size", [("config1", (320, 180)), ("moving_eye", (101, 67)), ("gpu_vs_cpu", (320, 200)),
                                       ("inner_given", (16, 9)), ("grid_a0.998_i89_d1", (48, 27))])
def test_route_b_frames_equal_the_software_renderer(ref, pair, name, size):
    desc, cam = cases.make(cases.CASES[name])
    s = ref.refSceneNew(C.byref(desc), C.byref(cam))
    try:
        print(f"route B {name} {size}", flush=True)
        sw, hip = frames(ref, pair, s, *size)
        assert sw.any()
        np.testing.assert_array_equal(hip, sw)
        w, h = size
        hdr = np.zeros((3, h, w), np.float32)
        assert ref.artfRendererReadHdr(pair[1], hdr[0].ctypes.data, hdr[1].ctypes.data, hdr[2].ctypes.data)
        want = np.zeros((3, h, w), np.float32)
        ref.refLens(s, w, h, 2, None, want[0].ctypes.data, want[1].ctypes.data, want[2].ctypes.data)
        np.testing.assert_array_equal(hdr.view(np.uint32), want.view(np.uint32))
    finally:
        ref.refSceneFree(s)


def test_route_b_overlay_and_background(ref, pair):
    """The overlay composed as the reference composes it, over a hole and over the background; the
    background with no hole, at full alpha and premultiplied at half."""
    desc, cam = cases.make(cases.CASES["config1"])
    s = ref.refSceneNew(C.byref(desc), C.byref(cam))
    w, h = 96, 70
    rng = np.random.default_rng(5)
    alpha = rng.integers(0, 256, (h, w), dtype=np.uint32)
    alpha[rng.random((h, w)) < .2] = 0
    alpha[rng.random((h, w)) < .2] = 255
    colour = (rng.random((h, w, 3)) * (alpha[..., None] + 1)).astype(np.uint32)  # premultiplied: no channel past alpha ...
    colour[alpha == 0] = rng.integers(0, 40, (int((alpha == 0).sum()), 3), dtype=np.uint32)  # ... but for glows of alpha 0
    overlay = (alpha << 24 | colour[..., 0] << 16 | colour[..., 1] << 8 | colour[..., 2]).astype(np.uint32)
    overlay[:8, :8] = 0
    overlay = np.ascontiguousarray(overlay[:, :80])  # narrower than the frame
    o = abi.Overlay(overlay.ctypes.data, overlay.shape[1], overlay.shape[0], 0, False, None, 0)
    try:
        for r in pair:
            ref.artfRendererSetOverlay(r, C.byref(o))
        sw, hip = frames(ref, pair, s, w, h)
        np.testing.assert_array_equal(hip, sw)
        ref.artfSceneSetBlackHole(s, None)
        for colour, a in (((0.0, 0.0, 1.0), 1.0), ((1.0, .5, .25), .5)):
            ref.artfSceneSetBackground(s, abi.Vec3(*colour))
            ref.artfSceneSetBackgroundAlpha(s, a)
            sw, hip = frames(ref, pair, s, w, h)
            np.testing.assert_array_equal(hip, sw)
        for r in pair:
            ref.artfRendererSetOverlay(r, None)
        sw, hip = frames(ref, pair, s, w, h)
        np.testing.assert_array_equal(hip, sw)
        assert (hip & 0xFFFFFF == (128 << 16 | 64 << 8 | 32)).all() and (hip >> 24 == 128).all()
    finally:
        for r in pair:
            ref.artfRendererSetOverlay(r, None)
        ref.refSceneFree(s)
