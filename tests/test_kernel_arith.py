"""Integer arithmetic the tracing kernel does in floating point, checked exhaustively on the CPU.

The kernel places a lane's pixel and ray (lane / side^2, pixel / tile width, ray / side, the shutter's
(5 ray) mod side^2) by smallQuotient: trunc((n + 1/2) x rcp(d)) for 0 <= n < 4096 and 1 <= d <= 64, with
the hardware's reciprocal, which is within 1 ulp of 1/d. Every such n and d, with the reciprocal taken 1 ulp
low, exact and 1 ulp high, must give the integer quotient."""
from __future__ import annotations

import numpy as np


def test_small_quotient_is_exact():
    n = np.arange(4096, dtype=np.float32)
    want_n = np.arange(4096)
    for d in range(1, 65):
        r = np.float32(1.0) / np.float32(d)
        for rcp in (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(2))):
            got = np.trunc(((n + np.float32(.5)) * rcp).astype(np.float32)).astype(np.int64)
            np.testing.assert_array_equal(got, want_n // d, err_msg=f"d {d}, rcp {rcp!r}")
