"""The headless drivers against the oracle.

artful-blackhole (tools/blackhole.c's loop on Linux) flies its orbit for a few fixed-step frames, and
artful-render (tools/render.c's black-hole scene) draws one picture after timed frames, and a sequence.
Each writes, beside its pictures, the hole's description and the camera that made them (--describe): the
eye's place, its orbital speed and the time are checked against the drivers' closed forms
(tools/blackhole.c:35-44, tools/render.c:336-360, with the orbital speed of a circular Kerr orbit
against the local static observers), and each picture must equal the oracle's render of that scene,
replayed on the host (the hole set at time 0, then moved on to the frame's time, as the drivers do), byte
for byte. The disc on the host or on the device gives the same frames, in either driver."""
from __future__ import annotations

import math
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from artful_amd import abi

pytestmark = pytest.mark.gpu

PKG = Path(__file__).resolve().parent.parent / "artful_amd"
BLACKHOLE, RENDER = PKG / "artful-blackhole", PKG / "artful-render"


def run(*cmd):
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def described(path):
    out = {}
    for line in Path(path).read_text().splitlines():
        _, k, _, d, _, c = line.split()
        out[int(k)] = (abi.BlackHoleDesc.from_buffer_copy(bytes.fromhex(d)), abi.Camera.from_buffer_copy(bytes.fromhex(c)))
    return out


def read_image(path):
    """RGB bytes (h, w, 3) of a P6 PPM or of a PNG of 8-bit RGBA rows with filter 0."""
    data = Path(path).read_bytes()
    if data.startswith(b"P6"):
        head = data.split(b"\n", 3)
        w, h = map(int, head[1].split())
        return np.frombuffer(head[3], np.uint8).reshape(h, w, 3)
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, idat, w = 8, b"", 0
    while at < len(data):
        n, = struct.unpack(">I", data[at:at + 4])
        kind, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        assert zlib.crc32(kind + body) == struct.unpack(">I", data[at + 8 + n:at + 12 + n])[0]
        if kind == b"IHDR":
            w, h, depth, colour = struct.unpack(">IIBB", body[:10])
            assert (depth, colour) == (8, 6)
        elif kind == b"IDAT":
            idat += body
        at += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert not rows[:, 0].any()
    rgba = rows[:, 1:].reshape(h, w, 4)
    assert (rgba[..., 3] == 255).all()
    return rgba[..., :3]


def oracle_rgb(native, oracle, desc, cam, w, h):
    """The scene a driver drew, replayed on the host: the hole set at time 0, then at the frame's time."""
    s = native.Scene()
    start = abi.BlackHoleDesc.from_buffer_copy(bytes(desc))
    start.time = 0.0
    s.set_hole(start)
    s.set_hole(desc)
    s.set_camera(cam)
    f, n = s.frame(w, h)
    _, _, bgra = oracle.render(f, s.map(), s.fluid()[0], n)
    s.close()
    return np.stack([(bgra >> 16) & 255, (bgra >> 8) & 255, bgra & 255], -1).astype(np.uint8)


def orbit_speed(a, r):
    """A circular equatorial Kerr orbit's speed against the static observers' frame-dragged rest there:
    (Omega - omega) varpi / alpha, Omega = 1 / (r^1.5 + a), omega = 2 a r / A."""
    delta, big = r * r - 2 * r + a * a, (r * r + a * a) ** 2 - a * a * (r * r - 2 * r + a * a)
    v = (1 / (r ** 1.5 + a) - 2 * a * r / big) * (math.sqrt(big) / r) / (r * math.sqrt(delta) / math.sqrt(big))
    return min(v, .99)


def vec(v):
    return np.array([v.x, v.y, v.z], np.float64)


def test_blackhole_driver_frames_are_the_oracle_s(gpu, oracle, tmp_path):
    assert BLACKHOLE.exists(), "build the library first: python -m artful_amd.build"
    w, h = 160, 90
    for where, extra in (("host", ()), ("device", ("--disc-on-device",))):
        out = tmp_path / where
        out.mkdir()
        run(BLACKHOLE, "--size", w, h, "--frames", 6, "--step", .2, "--save", 5, out, "--report", out / "report.txt",
            "--describe", out / "described.txt", *extra)
        assert "6 frames" in (out / "report.txt").read_text() and f"disc on the {where}" in (out / "report.txt").read_text()
    frames = described(tmp_path / "host" / "described.txt")
    assert sorted(frames) == [0, 5]
    around = 0.0
    for k in range(6):
        around += .2 / (74.1 * math.sqrt(74.1) + .6)  # the eye keeps an orbit's pace
        if k not in frames:
            continue
        desc, cam = frames[k]
        print(f"blackhole driver frame {k}", flush=True)
        # the eye 74.1 away at tilt .06, moving along its circle at the gas's speed, aimed where the hole appears
        tilt, d = .06, 74.1
        eye = d * np.array([math.sin(around) * math.cos(tilt), math.sin(tilt), math.cos(around) * math.cos(tilt)])
        speed = orbit_speed(.6, d)
        velocity = speed * np.array([math.cos(around), 0, -math.sin(around)])
        np.testing.assert_allclose(vec(cam.position), eye, atol=2e-4)
        np.testing.assert_allclose(vec(desc.velocity), velocity, atol=2e-6)
        np.testing.assert_allclose(vec(cam.target), velocity * d / math.sqrt(1 - speed * speed), atol=2e-4)
        assert abs(desc.time - .2 * (k + 1)) < 1e-5 and abs(cam.fovY - .33) < 1e-7
        got = read_image(tmp_path / "host" / f"frame-{k:05d}.ppm")
        assert np.array_equal(got, read_image(tmp_path / "device" / f"frame-{k:05d}.ppm"))
        np.testing.assert_array_equal(got, oracle_rgb(gpu, oracle, desc, cam, w, h))


def test_render_driver_picture_and_timing(gpu, oracle, tmp_path):
    """artful-render's black hole: 2x2 rays, a shutter, the eye in orbit 30 degrees around, after three
    timed frames (time .6: three fluid steps); its report line as the reference prints it."""
    assert RENDER.exists(), "build the library first: python -m artful_amd.build"
    w, h = 192, 108
    text = run(RENDER, "--backend", "hip", "--scene", "blackhole", "--size", w, h, "--samples", 2, "--shutter", .5, "--orbit",
               "--around", 30, "--doppler", 1, "--frames", 3, "--out", tmp_path / "hole.png", "--describe", tmp_path / "d.txt")
    print(text, flush=True)
    assert "3 frames:" in text and "ms a frame" in text and "on the GPU the last frame" in text
    desc, cam = described(tmp_path / "d.txt")[0]
    tilt, around, d = math.radians(3.44), math.radians(30), 74.1
    flat = math.cos(tilt) * d
    np.testing.assert_allclose(vec(cam.position), [math.sin(around) * flat, math.sin(tilt) * d, math.cos(around) * flat], atol=2e-4)
    speed = orbit_speed(.6, d)
    np.testing.assert_allclose(vec(desc.velocity), [math.cos(around) * speed, 0, -math.sin(around) * speed], atol=2e-6)
    assert abs(desc.time - .6) < 1e-5 and abs(cam.fovY - math.radians(19)) < 1e-6 and desc.samples == 2
    np.testing.assert_array_equal(read_image(tmp_path / "hole.png"), oracle_rgb(gpu, oracle, desc, cam, w, h))
    # the disc's map and fluid on the GPU: the same picture after the same timed frames
    text = run(RENDER, "--backend", "hip", "--scene", "blackhole", "--size", w, h, "--samples", 2, "--shutter", .5, "--orbit",
               "--around", 30, "--doppler", 1, "--frames", 3, "--out", tmp_path / "hole_device.png", "--disc-on-device")
    assert "3 frames:" in text
    np.testing.assert_array_equal(read_image(tmp_path / "hole_device.png"), read_image(tmp_path / "hole.png"))


def test_render_driver_sequence(gpu, oracle, tmp_path):
    """A sequence of three pictures, the eye going smoothly from 74.1 at 3.44 degrees to 40 at 20 degrees,
    the gas flowing .9 between pictures: each picture the oracle's."""
    w, h = 96, 54
    (tmp_path / "seq").mkdir()
    run(RENDER, "--size", w, h, "--sequence", 3, "--step", .9, "--to-distance", 40, "--to-tilt", 20, "--out",
        str(tmp_path / "seq" / "%04d.ppm"), "--describe", tmp_path / "d.txt")
    frames = described(tmp_path / "d.txt")
    assert sorted(frames) == [0, 1, 2]
    for i, (desc, cam) in frames.items():
        x = i / 2
        e = x * x * (3 - 2 * x)
        d, tilt = 74.1 + (40 - 74.1) * e, math.radians(3.44 + (20 - 3.44) * e)
        np.testing.assert_allclose(vec(cam.position), [0, math.sin(tilt) * d, math.cos(tilt) * d], atol=2e-4)
        assert abs(desc.time - .9 * i) < 1e-5
        np.testing.assert_array_equal(read_image(tmp_path / "seq" / f"{i:04d}.ppm"), oracle_rgb(gpu, oracle, desc, cam, w, h))
