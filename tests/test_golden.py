"""The host C library and the oracle against golden vectors the reference's own code produced
(tests/golden/make_fixtures.py, run on the reference's library built from its sources).

Bit-exact throughout: the disc map, the fluid, the frame uniforms, the clamps, the last stable
orbit; and the oracle's FP32 HDR and BGRA frames for every case."""
from __future__ import annotations

import hashlib
import math

import numpy as np
import pytest

from artful_amd import abi
from tests.golden import cases


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


NAMES = list(cases.CASES)


def test_defaults(native, golden):
    d = native.defaults()
    for k, v in golden["defaults"].items():
        got = getattr(d, k)
        got = got.tolist() if hasattr(got, "tolist") else got
        assert got == pytest.approx(v, abs=0), k


def test_isco(native, golden):
    for a, v in golden["isco"].items():
        assert np.float32(native.lib().artfHoleIsco(float(a))) == np.float32(v), a


def test_orbit_speed(native, golden):
    for key, v in golden["orbit_speed"].items():
        a, r = map(float, key.split(","))
        assert np.float32(native.lib().artfBlackHoleOrbitSpeed(a, r)) == np.float32(v), key


@pytest.mark.parametrize("name", NAMES)
def test_scene_data(name, scene_of, golden):
    """Map (gas and per-ring tables), fluid field and its moment, clamped description."""
    g = golden["scenes"][name]
    s = scene_of(name)
    m = s.map()
    assert digest(m[:abi.ROWS_AT]) == g["map_gas"]
    assert digest(m[abi.ROWS_AT:]) == g["map_tables"]
    fl, _, t = s.fluid()
    assert digest(fl) == g["fluid"]
    assert t == g["fluid_time"]
    d = s.desc()
    for k, v in g["clamped"].items():
        assert getattr(d, k) == pytest.approx(v, abs=0), k


@pytest.mark.parametrize("name", NAMES)
def test_frames(name, scene_of, golden):
    """The 368 bytes of artfHoleFrame and the bloom's size, at every size of the case."""
    g = golden["scenes"][name]
    s = scene_of(name)
    for size, want in g["frames"].items():
        w, h = map(int, size.split("x"))
        f, n = s.frame(w, h)
        assert bytes(f).hex() == want["bytes"], size
        assert n == want["bloom_pixels"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_hdr_and_bgra(name, scene_of, golden, oracle):
    """The oracle's FP32 trace equals the reference's lens kernels bit for bit, and its bloom and tone
    map give the reference software renderer's BGRA frame."""
    g = golden["scenes"][name]
    s = scene_of(name)
    m, (fl, _, _) = s.map(), s.fluid()
    for size in g["lens"]:
        w, h = map(int, size.split("x"))
        f, n = s.frame(w, h)
        hdr, _, bgra = oracle.render(f, m, fl, n)
        assert digest(hdr) == g["lens"][size], f"HDR {size}"
        assert digest(bgra) == g["bgra"][size], f"BGRA {size}"


def test_stored_hdr(scene_of, golden, oracle):
    """Where a whole frame is stored, compare values too (a readable failure)."""
    from pathlib import Path
    arrays = np.load(Path(__file__).parent / "golden" / "hdr_small.npz")
    for key in arrays.files:
        name, size, _ = key.rsplit("_", 2)
        w, h = map(int, size.split("x"))
        s = scene_of(name)
        f, _ = s.frame(w, h)
        hdr = oracle.trace(f, s.map(), s.fluid()[0])
        np.testing.assert_array_equal(hdr.view(np.uint32), arrays[key].view(np.uint32))


def test_advance(native, golden, oracle):
    """Time moves the gas (fluid advanced, map tables unchanged), as the reference does."""
    g = golden["scenes"]["advance"]
    desc, cam = cases.make(cases.CASES["advance"])
    s = native.Scene(desc, cam)
    version = s.map_version()
    desc.time = g["advanced"]["time"]
    s.set_hole(desc)
    assert s.map_version() == version
    fl, _, t = s.fluid()
    assert digest(fl) == g["advanced"]["fluid"]
    assert t == g["advanced"]["fluid_time"]
    assert digest(s.map()[abi.ROWS_AT:]) == g["advanced"]["map_tables"]
    w, h = cases.CASES["advance"]["sizes"][0]
    f, _ = s.frame(w, h)
    assert digest(oracle.trace(f, s.map(), fl)) == g["advanced"]["lens"]


def test_fluid_alone(native, golden):
    L = native.lib()
    for key, hashes in golden["fluid_alone"].items():
        inner, outer, spin = map(float, key.split(","))
        f = L.artfFluidNew(inner, outer, spin, 0.0)
        try:
            for t, want in hashes:
                L.artfFluidAdvance(f, float(t))
                field = np.ctypeslib.as_array(L.artfFluidField(f), shape=(abi.FLUID_WIDTH * abi.FLUID_HEIGHT,))
                assert digest(field) == want, t
        finally:
            L.artfFluidFree(f)


def test_fluid_same_however_cut(native):
    """tests.c:4180-4197: the same bytes however the time is cut, and time that does not reach a step
    changes nothing."""
    L = native.lib()
    a = L.artfFluidNew(3.83, 34.0, .6, 0.0)
    b = L.artfFluidNew(3.83, 34.0, .6, 0.0)
    try:
        n = abi.FLUID_WIDTH * abi.FLUID_HEIGHT
        before = np.ctypeslib.as_array(L.artfFluidField(a), shape=(n,)).copy()
        assert L.artfFluidAdvance(a, 2.0)
        assert L.artfFluidAdvance(b, 1.0) and L.artfFluidAdvance(b, 2.0)
        fa = np.ctypeslib.as_array(L.artfFluidField(a), shape=(n,))
        fb = np.ctypeslib.as_array(L.artfFluidField(b), shape=(n,))
        assert not np.array_equal(before, fa)
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        assert not L.artfFluidAdvance(a, L.artfFluidTime(a) + .1)
    finally:
        L.artfFluidFree(a)
        L.artfFluidFree(b)


def test_fluid_rings_turn_at_kerr_speed(native):
    """tests.c:4199-4222: each ring shifts by its own Kerr orbit, within 3 cells."""
    L = native.lib()
    inner, outer, spin = 3.83, 34.0, .6
    W, H = abi.FLUID_WIDTH, abi.FLUID_HEIGHT
    f = L.artfFluidNew(inner, outer, spin, 0.0)
    try:
        L.artfFluidAdvance(f, 2.0)
        before = np.ctypeslib.as_array(L.artfFluidField(f), shape=(H, W)).copy()
        start = L.artfFluidTime(f)
        assert L.artfFluidAdvance(f, start + 2.05)
        passed = L.artfFluidTime(f) - start
        now = np.ctypeslib.as_array(L.artfFluidField(f), shape=(H, W))
        for j in (90, 140):
            rho = j / (H - 1)
            r = inner / (1 - rho * (1 - inner / outer))
            expected = passed / (r * math.sqrt(r) + spin) / 6.2831853 * W
            gaps = [np.sum((before[j].astype(np.float64) - np.roll(now[j], -d)) ** 2) for d in range(W)]
            assert abs(int(np.argmin(gaps)) - expected) < 3.0, j
    finally:
        L.artfFluidFree(f)
