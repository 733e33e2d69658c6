"""One frame in one call (artfHipDrawFrame) and the bloom pyramid, against the oracle, bit for bit.

artfHipDrawFrame stands for artfHipTrace (which sums each pixel's rays in its lanes), artfHipBloom and
artfHipToneMap in a row (the reference's drawHole and lensFill, src/software/software.c:409-427,
lens.inc:355-440): its HDR, its
bloom pyramid and its BGRA must equal the oracle's for the same frame, and the step counts it leaves in
host memory must be the sums of the counters artfHipTrace fills for that frame, which it leaves zero.
The pyramid's small end is built by one workgroup in LDS: the sizes below put its first level at 0,
1, 2, 3 and 4, and leave no two levels small enough together, or one level alone."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from artful_amd import abi

pytestmark = pytest.mark.gpu

SLOTS = 1024  # ARTF_HIP_COUNTER_SLOTS


class Device:
    """Device buffers for one test, freed together."""

    def __init__(self, L):
        self.L, self.held = L, []

    def alloc(self, nbytes):
        p = self.L.artfHipMalloc(max(int(nbytes), 4))
        assert p, "device memory"
        self.held.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.artfHipUpload(p, a.ctypes.data, a.nbytes, None) == 0
        return p

    def planes(self, n, value=None):
        if value is None:
            return [self.alloc(n * 4) for _ in range(3)]
        return [self.upload(np.ascontiguousarray(value[c], np.float32)) for c in range(3)]

    def read(self, p, n, dtype=np.float32):
        out = np.zeros(n, dtype)
        assert self.L.artfHipDownload(out.ctypes.data, p, out.nbytes, None) == 0
        return out

    def free(self):
        for p in self.held:
            self.L.artfHipFree(p)


@pytest.fixture
def dev(gpu):
    d = Device(gpu.lib())
    yield d
    d.free()


def small_first(f):
    """Where the one-workgroup pass starts: the first level k whose levels k..last fit 32768 floats of
    LDS and whose levels k+1..last hold at most 1024 texels; the last level when none does."""
    last = f.size[3] - 1
    end = f.bloom[last][0] + f.bloom[last][1] * f.bloom[last][2]
    for k in range(last):
        if (end - f.bloom[k][0]) * 3 <= 32768 and end - f.bloom[k + 1][0] <= 1024:
            return k
    return last


PYRAMIDS = [(48, 27, 8, 0), (320, 200, 8, 2), (640, 400, 8, 3), (1920, 1080, 8, 4), (3840, 2160, 8, 5), (1920, 1080, 3, 2),
            (1920, 1080, 2, 1), (1920, 1080, 1, 0), (333, 77, 8, 1), (1001, 3, 8, 0), (1, 1, 8, 0), (5, 3, 4, 0)]


@pytest.mark.parametrize("w,h,levels,first", PYRAMIDS)
def test_bloom_and_tone_map_equal_oracle(w, h, levels, first, gpu, dev, scene_of, oracle):
    """artfHipBloom and artfHipToneMap on HDR with light on both sides of the bloom's threshold and past
    its ceiling, against oracleBloom and oracleToneMap."""
    L = gpu.lib()
    f, bloom_pixels = scene_of("config1").frame(w, h, levels)
    print(f"pyramid {w}x{h} levels {f.size[3]} small end from {small_first(f)}", flush=True)
    if first is not None:
        assert small_first(f) == first
    rng = np.random.default_rng(w * 7 + h * 3 + levels)
    hdr = (rng.exponential(2.0, (3, w * h)) * (rng.random((3, w * h)) < .3) * 30).astype(np.float32)
    d_hdr, d_bloom = dev.planes(w * h, hdr), dev.planes(bloom_pixels)
    d_bgra = dev.alloc(w * h * 4)
    P3 = C.c_void_p * 3
    assert L.artfHipBloom(C.byref(f), P3(*d_hdr), P3(*d_bloom), None) == 0
    assert L.artfHipToneMap(C.byref(f), P3(*d_hdr), P3(*d_bloom), d_bgra, None) == 0
    assert L.artfHipStreamSync(None) == 0
    got = np.stack([dev.read(d_bloom[c], bloom_pixels) for c in range(3)])
    want = oracle.bloom(f, hdr.reshape(3, h, w), bloom_pixels)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    bgra = dev.read(d_bgra, w * h, np.uint32).reshape(h, w)
    np.testing.assert_array_equal(bgra, oracle.tone_map(f, hdr.reshape(3, h, w), want))


FRAMES = [("config1", 320, 180), ("config2_small", 192, 108), ("moving_eye", 101, 67), ("inner_given", 16, 9), ("gpu_vs_cpu", 320, 200),
          ("moving_eye", 1, 1)]


def frame_args(L, dev, s, w, h, f, bloom_pixels):
    a = abi.HipFrameArgs()
    a.trace.frame = f
    a.trace.map, a.trace.fluid = dev.upload(s.map()), dev.upload(s.fluid()[0])
    hdr = dev.planes(w * h)
    for c, p in enumerate(hdr):
        a.trace.out[c] = p
    for c, (ph, pb) in enumerate(zip(hdr, dev.planes(bloom_pixels))):
        a.hdr[c], a.bloom[c] = ph, pb
    a.bgra = dev.alloc(w * h * 4)
    a.trace.counters = dev.upload(np.zeros(2 * SLOTS, np.uint64))
    return a


@pytest.mark.parametrize("streams", [1, 2])
@pytest.mark.parametrize("name,w,h", FRAMES)
def test_frame_equals_oracle(name, w, h, streams, gpu, dev, scene_of, oracle):
    """Two frames in a row through artfHipDrawFrame, on one stream or with the post passes handed to a
    second: HDR, pyramid and BGRA equal the oracle's render, the reported counts equal what artfHipTrace
    counts for the frame, the counters are left zero, and the stage stamps are in order."""
    L = gpu.lib()
    s = scene_of(name)
    f, bloom_pixels = s.frame(w, h)
    print(f"frame {name} {w}x{h} side {f.size[2]} levels {f.size[3]} streams {streams}", flush=True)
    a = frame_args(L, dev, s, w, h, f, bloom_pixels)
    report = L.artfHipHostMalloc(8 * abi.REPORT_WORDS)
    post, handoff, done = (L.artfHipStreamCreate(), L.artfHipEventCreateOrder(), L.artfHipEventCreateOrder()) if streams == 2 else (None,) * 3
    try:
        words = (C.c_uint64 * abi.REPORT_WORDS).from_address(report)
        a.report = report
        a.postStream, a.handoff, a.done = post, handoff, done
        counts = []
        for _ in range(2):
            for i in range(abi.REPORT_WORDS):
                words[i] = 0
            assert L.artfHipDrawFrame(C.byref(a), None) == 0
            assert L.artfHipStreamSync(None) == 0
            if post:
                assert L.artfHipStreamSync(post) == 0 and L.artfHipEventDone(done) == 1
            counts.append((words[0], words[1]))
            stamps = list(words[2:])
            print("clock stamps", [x - stamps[0] for x in stamps])
            assert 0 < stamps[0] <= stamps[1] <= stamps[2] <= stamps[3] <= stamps[4]
        assert counts[0] == counts[1]
        assert not dev.read(a.trace.counters, 2 * SLOTS, np.uint64).any(), "counters left zero"

        hdr = np.stack([dev.read(a.hdr[c], w * h) for c in range(3)]).reshape(3, h, w)
        pyramid = np.stack([dev.read(a.bloom[c], bloom_pixels) for c in range(3)])
        bgra = dev.read(a.bgra, w * h, np.uint32).reshape(h, w)
        want_hdr, want_bloom, want_bgra = oracle.render(f, s.map(), s.fluid()[0], bloom_pixels)
        np.testing.assert_array_equal(hdr.view(np.uint32), want_hdr.view(np.uint32))
        np.testing.assert_array_equal(pyramid.view(np.uint32), want_bloom.view(np.uint32))
        np.testing.assert_array_equal(bgra, want_bgra)

        # the same frame's counts from artfHipTrace alone
        t = abi.HipTraceArgs.from_buffer_copy(a.trace)
        assert L.artfHipTrace(C.byref(t), None) == 0
        assert L.artfHipStreamSync(None) == 0
        slots = dev.read(a.trace.counters, 2 * SLOTS, np.uint64)
        assert (int(slots[:SLOTS].sum()), int(slots[SLOTS:].sum())) == counts[0]
        assert 0 < counts[0][0] <= counts[0][1]
    finally:
        L.artfHipHostFree(report)
        if post:
            L.artfHipEventDestroy(handoff), L.artfHipEventDestroy(done), L.artfHipStreamDestroy(post)


@pytest.mark.parametrize("name,w,h", [("moving_eye", 101, 67), ("config1", 320, 180), ("inner_given", 16, 9)])
def test_renderer_frames_back_to_back(name, w, h, gpu, scene_of):
    """The renderer queues frames without waiting, each frame's bloom and tone map running beside the next
    frame's trace on a second stream, frames taking turns with two sets of HDR and pyramid. Frames of a
    changing scene drawn back to back, read now and then, equal the same frames drawn one at a time by a
    fresh renderer, bit for bit."""
    from tests.golden import cases
    desc, cam = cases.make(cases.CASES[name])
    times = [desc.time + .7 * k for k in range(7)]

    def scene_at(t):
        d = abi.BlackHoleDesc.from_buffer_copy(desc)
        d.time = t
        return d

    scenes = [gpu.Scene(scene_at(t), cam) for t in times]
    alone = {}
    r1 = gpu.Renderer(w, h)
    for k in (1, 4, 6):
        assert r1.draw(scenes[k])
        st = r1.hole_stats()
        alone[k] = (r1.read_pixels(), r1.read_hdr(), (st.steps, st.laneSlots))
    r1.close()

    r2 = gpu.Renderer(w, h)
    for k, sc in enumerate(scenes):
        assert r2.draw(sc)
        if k in alone:
            print(f"frame {k} at time {times[k]}", flush=True)
            px, hdr, counts = alone[k]
            np.testing.assert_array_equal(r2.read_hdr().view(np.uint32), hdr.view(np.uint32))
            np.testing.assert_array_equal(r2.read_pixels(), px)
            st = r2.hole_stats()
            assert (st.steps, st.laneSlots) == counts
            assert st.traceMs > 0 and st.bloomMs >= 0 and st.toneMs > 0
    r2.close()
    for sc in scenes:
        sc.close()


def test_uncounted_frames_are_the_same_frames(gpu):
    """ArtfHoleTracing.uncounted: frames queued back to back without step counts give the same pixels and
    HDR as counted ones, bit for bit, with steps and lane-slots 0; counted again, the counts come back
    (and the counted frame after uncounted ones counts only its own steps). The same for a batch."""
    from tests.golden import cases
    desc, cam = cases.make(cases.CASES["config1"])
    s = gpu.Scene(desc, cam)
    r = gpu.Renderer(160, 90)
    r.set_tracing("fp32")
    assert r.draw(s)
    want = (r.read_pixels(), r.read_hdr())
    st = r.hole_stats()
    counts = (st.steps, st.laneSlots)
    assert counts[0] > 0 and counts[0] <= counts[1]
    r.set_tracing("fp32", count_steps=False)
    for _ in range(3):
        assert r.draw(s)
    st = r.hole_stats()
    assert (st.steps, st.laneSlots) == (0, 0) and st.frames == 3
    np.testing.assert_array_equal(r.read_hdr().view(np.uint32), want[1].view(np.uint32))
    np.testing.assert_array_equal(r.read_pixels(), want[0])
    r.set_tracing("fp32")
    assert r.draw(s)
    st = r.hole_stats()
    assert (st.steps, st.laneSlots) == counts
    r.close()
    s.close()
    from artful_amd.workloads import uneven_views
    views, _ = uneven_views(count=3, seed=4, sizes=(64, 96))
    images = {}
    for count in (True, False):
        b = gpu.HoleBatch(views, gpu.tracing(precision="fp32", disc="painted", count_steps=count))
        b.draw()
        bs = b.stats()
        images[count] = [b.read_hdr(v) for v in range(len(views))]
        assert (bs.steps > 0) == count and (bs.laneSlots > 0) == count
        b.close()
    for a, c in zip(images[True], images[False]):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))


def test_renderer_resizes_and_precisions_between_queued_frames(gpu, scene_of):
    """Frames queued back to back across resizes (buffers replaced while frames may be in flight) and
    precision changes: each frame read back equals a fresh renderer's frame of the same size and
    tracing, bit for bit."""
    s = scene_of("moving_eye")
    plan = [((101, 67), "fp32"), ((320, 180), "fp32"), ((64, 36), "mixed"), ((320, 180), "fp32"), ((3, 2), "fp32_fast"), ((101, 67), "fp64")]
    want = []
    for (w, h), p in plan:
        r = gpu.Renderer(w, h)
        r.set_tracing(p)
        assert r.draw(s)
        want.append((r.read_pixels(), r.read_hdr()))
        r.close()
    r = gpu.Renderer(8, 8)
    for k, ((w, h), p) in enumerate(plan):
        r.resize(w, h)
        r.set_tracing(p)
        for _ in range(3):  # queued, not waited for
            assert r.draw(s)
        print(f"step {k}: {w}x{h} {p}", flush=True)
        np.testing.assert_array_equal(r.read_pixels(), want[k][0])
        np.testing.assert_array_equal(r.read_hdr().view(np.uint32), want[k][1].view(np.uint32))
        for _ in range(2):  # still in flight when the next step resizes
            assert r.draw(s)
    r.close()


def test_renderer_hole_removed_and_restored_between_queued_frames(gpu, native, scene_of):
    """Queued frames, then the hole taken out of the scene (the frame is the background), then put back:
    the frame after equals a fresh renderer's (renderer.h's scene contract, scene.h:497-500)."""
    from tests.golden import cases
    desc, cam = cases.make(cases.CASES["config1"])
    s = native.Scene(desc, cam)
    r0 = gpu.Renderer(64, 36)
    assert r0.draw(s)
    want = r0.read_pixels()
    r0.close()
    r = gpu.Renderer(64, 36)
    for _ in range(3):
        assert r.draw(s)
    s.set_hole(None)
    assert r.draw(s)
    assert (r.read_pixels() == 0xFF0D0F14).all()
    s.set_hole(desc)
    for _ in range(2):
        assert r.draw(s)
    np.testing.assert_array_equal(r.read_pixels(), want)
    r.close()
    s.close()


@pytest.mark.parametrize("side", range(1, 9))
def test_trace_every_side_equals_oracle(side, gpu, dev, native, oracle):
    """artfHipTrace at 1..8 rays a side (a wave holds 64 / side^2 pixels: 8 x 8, 4 x 4, 7 x 1, 2 x 2, 2 x 1
    and one pixel, with lanes left over at 3, 5, 6 and 7), on a frame that is no multiple of any of those
    tiles: each pixel's mean of its rays equals the oracle's, bit for bit, with a shutter (each ray its
    own moment) and a moving eye; and the step counts are the oracle's, and the lane-slots (each wave's
    longest ray times its rays) those the oracle's steps give."""
    from tests.golden import cases
    case = dict(cases.CASES["moving_eye"])
    case["desc"] = dict(case["desc"], samples=side)
    s = native.Scene(*cases.make(case))
    w, h = 23, 13
    f, _ = s.frame(w, h)
    assert f.size[2] == side
    L = gpu.lib()
    a = abi.HipTraceArgs()
    a.frame = f
    a.map, a.fluid = dev.upload(s.map()), dev.upload(s.fluid()[0])
    out = dev.planes(w * h)
    for c, p in enumerate(out):
        a.out[c] = p
    a.counters = dev.upload(np.zeros(2 * SLOTS, np.uint64))
    print(f"trace side {side}", flush=True)
    assert L.artfHipTrace(C.byref(a), None) == 0 and L.artfHipStreamSync(None) == 0
    got = np.stack([dev.read(out[c], w * h) for c in range(3)]).reshape(3, h, w)
    want, steps = oracle.trace(f, s.map(), s.fluid()[0], 32, steps=True)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    slots = dev.read(a.counters, 2 * SLOTS, np.uint64)
    assert int(slots[:SLOTS].sum()) == int(steps.astype(np.int64).sum())
    # lane-slots: each wave's longest ray times its rays, a wave being tw x th pixels' rays
    tw, th = {1: (8, 8), 2: (4, 4), 3: (7, 1), 4: (2, 2), 5: (2, 1)}.get(side, (1, 1))
    want_slots = 0
    for y in range(0, h, th):
        for x in range(0, w, tw):
            block = steps[:, y:y + th, x:x + tw].astype(np.int64)
            want_slots += int(block.max()) * block.size
    assert int(slots[SLOTS:].sum()) == want_slots
    s.close()
