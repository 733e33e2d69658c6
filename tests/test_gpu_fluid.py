"""The disc fluid stepped on the GPU (include/artful/device_disc.h) against the reference's fluid.

The fluid is chaotic: one wrong bit anywhere grows into a different flow within a few steps. So the
GPU's field must equal the reference's digests bit for bit (the standalone fluid at five moments, each
golden scene's fluid at its moment, the advanced scene's after a catch-up), and the host fluid's bytes
over a long run of single steps and jumps."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from artful_amd import abi
from tests.golden import cases

pytestmark = pytest.mark.gpu


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_fluid_alone_equals_reference(gpu, golden):
    for key, hashes in golden["fluid_alone"].items():
        inner, outer, spin = map(float, key.split(","))
        f = gpu.DeviceFluid(inner, outer, spin, 0.0)
        for t, want in hashes:
            f.advance(float(t))
            assert digest(f.field()) == want, (key, t)
        f.close()


@pytest.mark.parametrize("name", list(cases.CASES))
def test_scene_fluid_equals_reference(name, gpu, golden):
    g = golden["scenes"][name]
    case = cases.CASES[name]
    desc, _ = cases.make(case)
    c = g["clamped"]  # the reference's clamped description: the inner edge and spin the fluid is made for
    f = gpu.DeviceFluid(c["discInner"], c["discOuter"], c["spin"], float(desc.time))
    assert digest(f.field()) == g["fluid"] and f.time() == g["fluid_time"]
    if case.get("advance"):
        assert f.advance(case["advance"])
        assert digest(f.field()) == g["advanced"]["fluid"] and f.time() == g["advanced"]["fluid_time"]
    f.close()


@pytest.mark.parametrize("disc", [(6.0, 20.0, 0.0), (1.237, 30.0, .998), (5.0, 12.0, -.5)])
def test_long_run_equals_host(gpu, disc):
    """60 single steps, small and idle advances, a jump past the catch-up limit, a clock reset."""
    L = gpu.lib()
    inner, outer, spin = disc
    host = L.artfFluidNew(inner, outer, spin, 3.0)
    dev = gpu.DeviceFluid(inner, outer, spin, 3.0)
    times = [3.0 + .2 * k for k in range(1, 61)] + [15.05, 15.1, 40.0, 41.3, 20.0, 20.4]
    try:
        for t in times:
            a = L.artfFluidAdvance(host, t)
            b = dev.advance(t)
            want = np.ctypeslib.as_array(L.artfFluidField(host), shape=(abi.FLUID_WIDTH * abi.FLUID_HEIGHT,))
            assert a == b and L.artfFluidTime(host) == dev.time(), t
            assert np.array_equal(dev.field().view(np.uint32), want.view(np.uint32)), t
    finally:
        L.artfFluidFree(host)
        dev.close()


def test_renderer_with_the_disc_on_the_device(gpu):
    """A scene whose map is painted and whose fluid steps on the GPU draws the host scene's frames bit
    for bit, through an animation (a step a frame, as the reference's drivers advance time), a jump and
    a new disc."""
    from tests.golden import cases
    desc, cam = cases.make(cases.CASES["advance"])
    host, dev = gpu.Scene(desc, cam), gpu.Scene(cam=cam)
    dev.set_disc_on_device(0)
    dev.set_hole(desc)
    r = gpu.Renderer(96, 60)
    try:
        for k, t in enumerate([0.0, .2, .4, .6, .8, 1.0, 7.3, 7.5]):
            desc.time = t
            if k == 7:
                desc.spin, desc.discOuter = .9, 25.0
            host.set_hole(desc)
            dev.set_hole(desc)
            assert dev.fluid()[2] == host.fluid()[2], t
            assert r.draw(host)
            want, want_px = r.read_hdr(), r.read_pixels()
            assert r.draw(dev)
            assert np.array_equal(r.read_hdr().view(np.uint32), want.view(np.uint32)), t
            assert np.array_equal(r.read_pixels(), want_px), t
    finally:
        r.close(), host.close(), dev.close()


def test_queued_frames_with_the_disc_on_the_device(gpu):
    """Frames of a scene whose fluid steps on the GPU, queued back to back with no wait between them, each
    on a renderer of its own (so each frame's HDR stays to be read afterwards): every fluid step writes the
    field buffer no queued frame reads, and every frame reads the field of its own moment, so each frame
    equals the host scene's frame at that moment bit for bit. Then ten renderers read one field, more
    than the fluid keeps read marks for, before the fluid steps twice and is read again."""
    desc, cam = cases.make(cases.CASES["advance"])
    host, dev = gpu.Scene(desc, cam), gpu.Scene(cam=cam)
    dev.set_disc_on_device(0)
    dev.set_hole(desc)
    times = [round(.2 * k, 1) for k in range(1, 11)]
    rs = [gpu.Renderer(96, 60) for _ in times]
    ref = gpu.Renderer(96, 60)
    try:
        for r in rs:  # buffers made and nothing left queued before the timed part
            assert r.draw(dev)
            r.read_hdr()
        for r, t in zip(rs, times):
            desc.time = t
            dev.set_hole(desc)
            assert r.draw(dev)
        got = [r.read_hdr() for r in rs]
        for k, t in enumerate(times):
            desc.time = t
            host.set_hole(desc)
            assert ref.draw(host)
            assert np.array_equal(got[k].view(np.uint32), ref.read_hdr().view(np.uint32)), t
        for r in rs:  # ten reads of the field of one moment
            assert r.draw(dev)
        for t in (times[-1] + .2, times[-1] + .4):
            desc.time = t
            dev.set_hole(desc)
        assert rs[0].draw(dev)
        want_same = ref.read_hdr()  # the host frame at times[-1]
        for r in rs[1:]:
            assert np.array_equal(r.read_hdr().view(np.uint32), want_same.view(np.uint32))
        for t in (times[-1] + .2, times[-1] + .4):
            desc.time = t
            host.set_hole(desc)
        assert ref.draw(host)
        assert np.array_equal(rs[0].read_hdr().view(np.uint32), ref.read_hdr().view(np.uint32))
    finally:
        for r in rs:
            r.close()
        ref.close(), host.close(), dev.close()
