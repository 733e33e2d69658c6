"""artfHipDrawFrameStatic, the allocation-free artfHipDrawFrame, against artfHipDrawFrame itself through
the C ABI: the same HDR, bloom pyramid, BGRA and reported step counts, bit for bit, on two calls in a row,
the second on a smaller frame than the first (the static counters, which the first call leaves zero, are
shared); and the refusals on the host, before any launch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from artful_amd import abi
from tests.golden import cases

pytestmark = pytest.mark.gpu

SLOTS = 1024  # ARTF_HIP_COUNTER_SLOTS
INVALID = 1   # hipErrorInvalidValue


class Buffers:
    def __init__(self, L):
        self.L, self.held = L, []

    def alloc(self, nbytes):
        p = self.L.artfHipMalloc(nbytes)
        assert p, "device memory"
        self.held.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.artfHipUpload(p, a.ctypes.data, a.nbytes, None) == 0
        return p

    def read(self, p, n, dtype=np.float32):
        out = np.zeros(n, dtype)
        assert self.L.artfHipDownload(out.ctypes.data, p, out.nbytes, None) == 0
        return out

    def free(self):
        for p in self.held:
            self.L.artfHipFree(p)


def frame_args(buf, s, w, h):
    f, bloom_pixels = s.frame(w, h)
    a = abi.HipFrameArgs()
    a.trace.frame = f
    a.trace.map, a.trace.fluid = buf.upload(s.map()), buf.upload(s.fluid()[0])
    for c in range(3):
        a.hdr[c] = a.trace.out[c] = buf.alloc(w * h * 4)
        a.bloom[c] = buf.alloc(bloom_pixels * 4)
    a.bgra = buf.alloc(w * h * 4)
    return a, bloom_pixels


def outputs(buf, a, w, h, bloom_pixels):
    hdr = np.stack([buf.read(a.hdr[c], w * h) for c in range(3)])
    pyramid = np.stack([buf.read(a.bloom[c], bloom_pixels) for c in range(3)])
    return hdr, pyramid, buf.read(a.bgra, w * h, np.uint32)


def test_static_frame_equals_draw_frame(gpu, native):
    L = gpu.lib()
    buf = Buffers(L)
    reports = [L.artfHipHostMalloc(8 * abi.REPORT_WORDS) for _ in range(2)]
    try:
        # a moving eye at 2 x 2 rays, then the CPU-sized view, smaller, at one ray a pixel
        for name, (w, h) in (("gpu_vs_cpu", (320, 200)), ("config1", (96, 54))):
            desc, cam = cases.make(cases.CASES[name])
            s = native.Scene(desc, cam)
            print(f"static frame {name} {w}x{h}", flush=True)
            got, want = frame_args(buf, s, w, h), frame_args(buf, s, w, h)
            a, bloom_pixels = got
            b, _ = want
            b.trace.counters = buf.upload(np.zeros(2 * SLOTS, np.uint64))
            words = [(C.c_uint64 * abi.REPORT_WORDS).from_address(r) for r in reports]
            a.report, b.report = reports
            a.trace.counters = None
            assert L.artfHipDrawFrameStatic(C.byref(a), None) == 0
            assert L.artfHipDrawFrame(C.byref(b), None) == 0
            assert L.artfHipStreamSync(None) == 0
            for x, y in zip(outputs(buf, a, w, h, bloom_pixels), outputs(buf, b, w, h, bloom_pixels)):
                np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
            assert (words[0][0], words[0][1]) == (words[1][0], words[1][1]) and words[0][0] > 0
            s.close()
    finally:
        for r in reports:
            L.artfHipHostFree(r)
        buf.free()


def test_static_frame_refusals(gpu, native):
    """Refused on the host: no arguments, a second stream (the static counters would be shared across
    streams), a frame taller than a grid, an output that is not the HDR; nothing launched."""
    L = gpu.lib()
    assert L.artfHipDrawFrameStatic(None, None) == INVALID
    buf = Buffers(L)
    try:
        desc, cam = cases.make(cases.CASES["config1"])
        s = native.Scene(desc, cam)
        a, _ = frame_args(buf, s, 16, 9)
        stream = L.artfHipStreamCreate()
        a.postStream, a.handoff = stream, stream
        assert L.artfHipDrawFrameStatic(C.byref(a), None) == INVALID
        a.postStream = a.handoff = None
        a.trace.frame.size[1] = 65536
        assert L.artfHipDrawFrameStatic(C.byref(a), None) == INVALID
        a.trace.frame.size[1] = 9
        a.trace.out[0] = a.bloom[0]
        assert L.artfHipDrawFrameStatic(C.byref(a), None) == INVALID
        L.artfHipStreamDestroy(stream)
        s.close()
    finally:
        buf.free()
