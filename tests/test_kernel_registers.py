"""The register budgets the frame path relies on, read from the kernels' own compiled metadata (hipcc
for gfx950, device code only, the library's flags): the exact FP32 tracer, and the batch's kernel with a
workgroup a tile, which traces with its code, in at most 96 VGPRs, so that a SIMD runs 5 of their waves, with at most two registers spilled (a value held across the step loop, stored
once as the ray starts and read once as it ends: 8 bytes of scratch a lane, none of it in the loop); and the post kernels a frame runs beside the next frame's
trace (bloom level 0, a level down, a level up, the tone map) in at most 32, the registers a SIMD then has
left, so that they start as soon as a trace wave ends. A compiler that gave any of them more would leave
the frames' post passes waiting for the trace to drain."""
from __future__ import annotations

import re
import shutil
import subprocess

import pytest

from artful_amd import build as B

pytestmark = pytest.mark.skipif(shutil.which(B.hipcc()) is None and not __import__("os").path.exists(B.hipcc()), reason="no hipcc")


def kernels(source: str, contract: str, more: list[str], tmp_path) -> dict[str, tuple[int, int]]:
    out = tmp_path / "k.s"
    cmd = [B.hipcc(), *B.hip_flags(contract), *more, "--offload-device-only", "-S", "-c", str(B.CSRC / source), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = {}
    for block in out.read_text().split("  - .")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block)
        vgprs = re.search(r"\.vgpr_count:\s+(\d+)", block)
        spills = re.search(r"\.vgpr_spill_count:\s+(\d+)", block)
        if name and vgprs and spills:
            meta[name.group(1)] = (int(vgprs.group(1)), int(spills.group(1)))
    return meta


def one(meta, part):
    found = [v for k, v in meta.items() if part in k]
    assert len(found) == 1, (part, list(meta))
    return found[0]


def test_exact_tracer_fits_five_waves(tmp_path):
    meta = kernels("hip/trace_exact.hip", "off", B.ILP, tmp_path)
    for part in ("traceKernelILi0ELb0E", "traceKernelILi0ELb1E"):  # the drag's quotient by IEEE division, and in range
        vgprs, spills = one(meta, part)
        assert vgprs <= 96 and spills <= 2, (part, vgprs, spills)
    # the batch's kernel with a workgroup a tile traces with the same code at the same budget
    for part in ("viewTilesKernelILi0ELb0E", "viewTilesKernelILi0ELb1E"):
        vgprs, spills = one(meta, part)
        assert vgprs <= 96 and spills <= 2, (part, vgprs, spills)


def test_frame_post_kernels_fit_beside_the_trace(tmp_path):
    meta = kernels("hip/post.hip", "off", [], tmp_path)
    for part in ("bloomFirstKernel", "bloomDownKernel", "bloomUpKernel", "toneKernel"):
        vgprs, spills = one(meta, part)
        assert vgprs <= 32 and spills == 0, (part, vgprs, spills)
