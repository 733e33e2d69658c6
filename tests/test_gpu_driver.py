"""artful-blackhole, the headless driver (tools/blackhole.c's loop on Linux): it flies the orbit for a
few fixed-step frames, writes PPM frames and its report, and gives the same frames with the disc's map
and fluid on the host or on the device."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent.parent / "artful_amd" / "artful-blackhole"


def fly(tmp_path, name, *extra):
    out = tmp_path / name
    out.mkdir()
    report = out / "report.txt"
    r = subprocess.run([str(DRIVER), "--size", "160", "90", "--frames", "6", "--step", ".2", "--save", "5", str(out),
                        "--report", str(report), *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out, report.read_text()


def test_driver_flies_and_reports(gpu, tmp_path):
    assert DRIVER.exists(), "build the library first: python -m artful_amd.build"
    host, text = fly(tmp_path, "host")
    dev, text_dev = fly(tmp_path, "device", "--disc-on-device")
    assert "6 frames" in text and "disc on the host" in text and "disc on the device" in text_dev
    frames = sorted(p.name for p in host.glob("frame-*.ppm"))
    assert frames == ["frame-00000.ppm", "frame-00005.ppm"]
    for name in frames:
        a, b = (host / name).read_bytes(), (dev / name).read_bytes()
        assert a.startswith(b"P6\n160 90\n255\n") and len(a) == len(b"P6\n160 90\n255\n") + 160 * 90 * 3
        assert a == b, name
    assert (host / frames[0]).read_bytes() != (host / frames[1]).read_bytes()  # the eye moved, the gas turned
