"""Arithmetic the tracing kernel does its own way, checked on the CPU.

The kernel places a lane's pixel and ray (lane / side^2, pixel / tile width, ray / side, the shutter's
(5 ray) mod side^2) by smallQuotient: trunc((n + 1/2) x rcp(d)) for 0 <= n < 4096 and 1 <= d <= 64, with
the hardware's reciprocal, which is within 1 ulp of 1/d. Every such n and d, with the reciprocal taken 1 ulp
low, exact and 1 ulp high, must give the integer quotient.

The steps' drag takes its quotient by quotInRange (tracer.h): the IEEE division's own steps without its
scaling and fix-up, and the sign of a zero put back. tests/arith/quot_in_range.c (built here with gcc)
runs those steps with C's fmaf over the range the kernel claims for them, with the reciprocal 1 ulp either
way, against IEEE division, bit for bit; and checks that the drag's divisor is exactly 1 wherever
|u| < 2^-26, where its dividend may be tiny."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest


def test_small_quotient_is_exact():
    n = np.arange(4096, dtype=np.float32)
    want_n = np.arange(4096)
    for d in range(1, 65):
        r = np.float32(1.0) / np.float32(d)
        for rcp in (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(2))):
            got = np.trunc(((n + np.float32(.5)) * rcp).astype(np.float32)).astype(np.int64)
            np.testing.assert_array_equal(got, want_n // d, err_msg=f"d {d}, rcp {rcp!r}")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_quot_in_range_is_ieee_division(tmp_path):
    src = Path(__file__).resolve().parent / "arith" / "quot_in_range.c"
    exe = tmp_path / "quot_in_range"
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", str(src), "-o", str(exe), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 wrong; 0 divisors not 1" in r.stdout, r.stdout
