"""The host C code under the sanitizers, with its allocations counted, and its bytes whatever the
thread count (the reference runs its tests under ASan and UBSan, .github/workflows/ci.yml:41-46, counts
the blocks alive with ARTFUL_COUNT_ALLOCATIONS, src/core/base.c:32-40 and tests.c:4560-4563, and holds a
frame's bytes equal for any thread count, tests.c:4317-4340).

The library's host sources (memory, pool, disc map, fluid, frame, scene) are built on their own with
gcc -fsanitize=address,undefined and ARTFUL_COUNT_ALLOCATIONS into tests/hygiene/runner.c, which makes
golden scenes at 1, 3 and 16 threads: no sanitizer report, no block left alive, and the map, fluid and
frame bytes equal at every thread count and to the reference's digests."""
from __future__ import annotations

import hashlib
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.golden import cases

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "artful_amd" / "csrc"
SOURCES = ["core/base.c", "core/pool.c", "scene/disc.c", "scene/fluid.c", "scene/hole.c", "scene/frame64.c", "scene/scene.c"]


def digest(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()[:16]


@pytest.fixture(scope="module")
def runner(tmp_path_factory):
    out = tmp_path_factory.mktemp("hygiene") / "runner"
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-ffp-contract=off", "-fno-math-errno", "-fno-trapping-math", "-DARTFUL_COUNT_ALLOCATIONS", "-Wall", "-Wno-parentheses", "-Werror=implicit-function-declaration",
           f"-I{ROOT / 'include'}", f"-I{CSRC}", str(ROOT / "tests" / "hygiene" / "runner.c"),
           *[str(CSRC / s) for s in SOURCES], "-lm", "-lpthread", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("name,times", [("config1", "0"), ("advance", "0,40")])
def test_host_code_clean_and_the_same_at_any_thread_count(runner, golden, tmp_path, name, times):
    case = cases.CASES[name]
    desc, cam = cases.make(case)
    w, h = case["sizes"][0]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(runner), str(tmp_path), bytes(desc).hex(), bytes(cam).hex(), str(w), str(h), times, "1", "3", "16"],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "allocations alive 0" in r.stdout, r.stdout
    g = golden["scenes"][name]
    for what in ("map", "fluid", "frame"):
        blobs = [(tmp_path / f"{what}-{t}.bin").read_bytes() for t in (1, 3, 16)]
        assert blobs[0] == blobs[1] == blobs[2], what
    m = np.frombuffer((tmp_path / "map-1.bin").read_bytes(), np.float32)
    from artful_amd import abi
    want = g["advanced"] if name == "advance" else g
    assert digest(m[:abi.ROWS_AT].tobytes()) == g["map_gas"]
    assert digest(m[abi.ROWS_AT:].tobytes()) == want["map_tables"]
    assert digest((tmp_path / "fluid-1.bin").read_bytes()) == want["fluid"]
    assert (tmp_path / "frame-1.bin").read_bytes().hex() == g["frames"][f"{w}x{h}"]["bytes"] or name == "advance"
