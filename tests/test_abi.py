"""The C-ABI library without a GPU: it loads, exports every function include/*.h declares, its structs
have the sizes the headers give, and its entry points reject bad arguments before launching anything."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import pytest

from artful_amd import abi

ROOT = Path(__file__).resolve().parent.parent


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
    assert C.sizeof(abi.HipTraceArgs) == 816 and C.sizeof(abi.HipFrameArgs) == 904
    assert C.sizeof(abi.HipView) == 800 and C.sizeof(abi.HipBatchArgs) == 104 and C.sizeof(abi.HoleBatchStats) == 64 and C.sizeof(abi.HoleSchedule) == 20
    assert abi.MAP_FLOATS == 2887424


LAYOUT_TYPES = {"ArtfRendererOptions": abi.RendererOptions, "ArtfSurface": abi.Surface, "ArtfFrameStats": abi.FrameStats,
                "ArtfOverlay": abi.Overlay, "ArtfVec3": abi.Vec3, "ArtfCamera": abi.Camera, "ArtfBlackHoleDesc": abi.BlackHoleDesc,
                "ArtfHoleFrame": abi.HoleFrame}


def test_public_layouts_are_the_reference_s(golden, tmp_path):
    """Every public struct a program compiled against the reference hands this library (options, stats,
    overlay, camera, hole description; and the frame's uniform block) has the reference's size and
    field offsets, and the enums their values: sizeof/offsetof of the reference's own headers
    (oracle/ref/ref_driver.c) against this library's headers, compiled here, and its ctypes mirrors."""
    want = golden["layouts"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "artful/renderer.h"', '#include "artful/hole.h"', 'int main(void) {']
    for ty, fields in want.items():
        if ty == "enums":
            continue
        for field in fields:
            expr = f"sizeof({ty})" if field == "sizeof" else f"offsetof({ty}, {field})"
            lines.append(f'    printf("{ty} {field} %zu\\n", {expr});')
    for name in want["enums"]:
        lines.append(f'    printf("enums {name} %d\\n", (int){name});')
    lines.append("    return 0;\n}")
    src = tmp_path / "layouts.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layouts"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got: dict = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        ty, field, value = line.split()
        got.setdefault(ty, {})[field] = int(value)
    assert got == want
    for ty, cls in LAYOUT_TYPES.items():
        assert C.sizeof(cls) == want[ty]["sizeof"], ty
        for field, at in want[ty].items():
            if field != "sizeof":
                assert getattr(cls, field).offset == at, (ty, field)


REFERENCE = Path("/root/reference")


@pytest.mark.skipif(not (REFERENCE / "src" / "render" / "rendererInternal.h").exists(), reason="the reference tree is not here")
def test_route_b_compiles_against_the_reference(tmp_path):
    """The HIP backend (hip_renderer.c) and the route-B scene port compile against the reference's own
    headers, the reference's include/ and src/ ahead of this repository's, with every warning on: the
    backend reads the reference's scene and fills the reference's vtable as they are."""
    csrc = ROOT / "artful_amd" / "csrc"
    for f in ("render/hip_renderer.c", "render/route_b/hip_scene_reference.c", "scene/frame64.c"):
        r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                            "-Wno-sign-compare", f"-I{REFERENCE / 'include'}", f"-I{REFERENCE / 'src'}", f"-I{ROOT / 'include'}",
                            f"-I{csrc}", "-M", str(csrc / f)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        deps = r.stdout
        assert str(REFERENCE / "include" / "artful" / "scene.h") in deps and str(ROOT / "include" / "artful" / "scene.h") not in deps
        if f == "render/hip_renderer.c":  # the reference's vtable and renderer API
            assert str(REFERENCE / "src" / "render" / "rendererInternal.h") in deps
            assert str(REFERENCE / "include" / "artful" / "renderer.h") in deps
            assert str(csrc / "render" / "rendererInternal.h") not in deps
        if "route_b" in f:  # the reference's scene as the Vulkan backend reads it
            assert str(REFERENCE / "src" / "scene" / "sceneInternal.h") in deps
            assert "scene_internal.h" not in deps
        r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                            "-Wno-sign-compare", f"-I{REFERENCE / 'include'}", f"-I{REFERENCE / 'src'}", f"-I{ROOT / 'include'}",
                            f"-I{csrc}", str(csrc / f)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


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
    the HDR as the trace's output with one ray a pixel, a handoff event with a second stream, whole frames
    only."""
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
