"""The C-ABI library without a GPU: it loads, exports every function include/*.h declares, its structs
have the sizes the headers give, and its entry points reject bad arguments before launching anything."""
from __future__ import annotations

import ctypes as C
import subprocess

import pytest

from artful_amd import abi


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_every_declared_symbol_is_exported(native):
    names = native.declared_symbols()
    assert len(names) > 50
    have = exported(native.LIB_PATH)
    missing = [n for n in names if n not in have]
    assert not missing, missing


def test_struct_sizes():
    assert C.sizeof(abi.BlackHoleDesc) == 92
    assert C.sizeof(abi.HoleFrame) == 368
    assert C.sizeof(abi.HoleFrame64) == 352
    assert C.sizeof(abi.HipTraceArgs) == 808 and C.sizeof(abi.HipFrameArgs) == 896
    assert C.sizeof(abi.HipView) == 776 and C.sizeof(abi.HipBatchArgs) == 96 and C.sizeof(abi.HoleBatchStats) == 64 and C.sizeof(abi.HoleSchedule) == 20
    assert abi.MAP_FLOATS == 2887424


def test_backend_names(native):
    L = native.lib()
    assert [L.artfBackendName(i).decode() for i in range(5)] == ["auto", "Vulkan", "software", "Metal", "HIP"]


def test_renderer_contract_without_device(native):
    """artfRendererNew returns NULL (and logs) when the backend asked for cannot run; asking for the
    reference's other backends by name also gives NULL here."""
    L = native.lib()
    if native.device_count() > 0:
        pytest.skip("a GPU is present")
    for backend in (abi.BACKEND_AUTO, abi.BACKEND_HIP, abi.BACKEND_SOFTWARE, abi.BACKEND_VULKAN):
        o = L.artfRendererDefaults(16, 16)
        o.backend = backend
        assert not L.artfRendererNew(C.byref(o))


def test_trace_rejects_bad_arguments(native):
    """Validation happens on the host, before any launch: NULL pointers, sizes out of range."""
    L = native.lib()
    a = abi.HipTraceArgs()
    assert L.artfHipTrace(C.byref(a), None) != 0          # empty frame
    a.frame.size[0], a.frame.size[1], a.frame.size[2], a.frame.size[3] = 16, 16, 1, 1
    assert L.artfHipTrace(C.byref(a), None) != 0          # no buffers
    a.frame.size[2] = 9
    assert L.artfHipTrace(C.byref(a), None) != 0          # side beyond 8
    a.frame.size[2] = 1
    a.map = a.fluid = 1
    a.out[0] = a.out[1] = a.out[2] = 1
    a.frame.size[0], a.frame.size[1] = 1, 65536
    assert L.artfHipTrace(C.byref(a), None) != 0          # taller than a grid can be
    a.frame.size[0], a.frame.size[1], a.frame.size[2] = 65536, 65535, 8
    assert L.artfHipTrace(C.byref(a), None) != 0          # 2^31 rays or more
    f = abi.HoleFrame()
    f.size[0], f.size[1], f.size[2], f.size[3] = 16, 16, 1, 1
    P3 = C.c_void_p * 3
    assert L.artfHipBloom(C.byref(f), P3(1, 1, 1), P3(1, 1, 1), None) != 0  # level table not artfHoleFrame's
    assert L.artfHipToneMap(C.byref(f), P3(1, 1, 1), P3(1, 1, 1), None, None) != 0


def test_frame_rejects_bad_arguments(native):
    """artfHipDrawFrame checks what the four calls it stands for check, and its own: the output planes,
    the HDR as the trace's output with one ray a pixel, counters under a report, whole frames only."""
    L = native.lib()
    INVALID = 1  # hipErrorInvalidValue: refused on the host, nothing launched
    assert L.artfHipDrawFrame(None, None) == INVALID
    a = abi.HipFrameArgs()
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID  # empty frame
    f = abi.HoleFrame()
    w, h = 16, 16
    f.size[0], f.size[1], f.size[2], f.size[3] = w, h, 2, 1
    f.bloom[0][0], f.bloom[0][1], f.bloom[0][2] = 0, 8, 8
    a.trace.frame = f
    a.trace.map = a.trace.fluid = 64
    for c in range(3):
        a.trace.out[c], a.hdr[c], a.bloom[c] = 64 * (c + 1), 64 * (c + 4), 64 * (c + 7)
    a.bgra = 64
    a.report = 64  # a report without counters
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID
    a.report = None
    a.postStream = 64  # a second stream without the handoff event
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID
    a.postStream = None
    a.trace.rowBegin = 4  # a band, not a frame
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID
    a.trace.rowBegin = 0
    a.bgra = None
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID
    a.bgra = 64
    a.trace.frame.bloom[0][1] = 7  # a level table not artfHoleFrame's
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID
    a.trace.frame.bloom[0][1] = 8
    a.trace.frame.size[2] = 1  # one ray a pixel: the trace must write the HDR planes
    assert L.artfHipDrawFrame(C.byref(a), None) == INVALID


def test_batch_rejects_bad_arguments(native):
    """The batch's kernel entries check their schedule and buffers on the host; artfHoleBatchNew checks
    its views and tracing before it looks for a device."""
    L = native.lib()
    a = abi.HipBatchArgs()
    assert L.artfHipTraceViews(C.byref(a), None) != 0 and L.artfHipBatchReset(C.byref(a), None) != 0
    a.views = a.firstTile = a.map = a.rings = a.queue = 64
    a.viewCount, a.tileBegin, a.tileEnd, a.precision = 1, 0, 1, 0
    for bad in [(0, 1, 1, 1), (1, 0, 1, 1), (1, 65, 1, 1), (1, 1, 0, 1), (1, 1, 65, 1), (1, 1, 1, 0), (1, 1, 1, 17)]:
        a.schedule = abi.HipSchedule(*bad)
        assert L.artfHipTraceViews(C.byref(a), None) != 0, bad
    a.schedule = abi.HipSchedule(1, 1, 1, 1)
    a.viewCount = 70000  # more views than a batch holds
    assert L.artfHipTraceViews(C.byref(a), None) != 0
    a.viewCount, a.tileBegin = 1, 1  # an empty slice
    assert L.artfHipTraceViews(C.byref(a), None) != 0
    a.tileBegin, a.precision = 0, 4
    assert L.artfHipTraceViews(C.byref(a), None) != 0

    desc, cam = native.hole_desc(), native.camera((0, 10, 50))
    views = (abi.HoleView * 2)()
    for v in views:
        v.hole, v.camera, v.width, v.height = desc, cam, 16, 16
    t = L.artfHoleTracingDefaults()
    assert not L.artfHoleBatchNew(views, 0, C.byref(t), 0)
    views[1].width = 0
    assert not L.artfHoleBatchNew(views, 2, C.byref(t), 0)
    views[1].width, t.precision = 16, 7
    assert not L.artfHoleBatchNew(views, 2, C.byref(t), 0)
    t.precision, t.maxSteps = 0, -1
    assert not L.artfHoleBatchNew(views, 2, C.byref(t), 0)


def test_scene_lifecycle(native):
    s = native.Scene()
    assert s.desc() is None
    d = native.hole_desc(spin=2.0, samples=20, discHeight=0.0, mass=0.0)
    s.set_hole(d)
    got = s.desc()
    assert got.spin == pytest.approx(.998) and got.samples == 8 and got.discHeight == pytest.approx(1e-3) and got.mass > 0
    v = s.map_version()
    s.set_hole(d)             # same disc and spin: no new map
    assert s.map_version() == v
    d.discOuter = 12.0        # another disc: new tables
    s.set_hole(d)
    assert s.map_version() == v + 1
    s.set_hole(None)
    assert s.desc() is None
    s.close()
