"""Writes tests/golden/steps.json: how many Nystrom steps each ray of each configuration takes, counted
by the oracle (the C restatement of lensTrace, oracle/lens_ray.inc, whose FP32 light equals the
reference's lens kernels bit for bit on the golden scenes), so that the GPU's step counters, and the
FLOPs bench.py derives from them, are scored on the same sums the CPU path takes.

For each configuration's reference-parity form (each view on a host scene of its own: map, rings and
fluid made by the host code the reference's digests pin):
- rays, and steps: the sum over rays of the steps each takes to its own exit (escape, horizon, opacity,
  or the step cap);
- histogram: [steps, rays] pairs, every step count that occurs;
- lane_slots (the renderer's configurations): what the tile tracer's counters report, the sum over its
  waves of the wave's longest ray times its rays, a wave holding the rays of a tile of whole pixels (8x8
  pixels at one ray a pixel, 4x4 at 2x2 rays, ...);
- views (the batch): the steps of each view.

Configurations: 1 the CPU-sized view (320x180, spin .9); 2 the 1080p workload (1920x1080, 2x2 rays,
spin .998); 3 the precision study (3840x2160, spins {0, .5, .9, .998} x inclinations {17, 45, 75, 89});
4 the 64 uneven views of seed 4 (up to 5000 steps); 5 the adjoint's view (512x512, analytic disc, q 2.5,
up to 2000 steps, escape at r > 100), at its hidden parameters.

    python tests/golden/make_steps.py            # all (a few minutes on 8 cores)
    python tests/golden/make_steps.py config1    # some
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from artful_amd import native  # noqa: E402
from artful_amd.workloads import uneven_views, view_of  # noqa: E402
from oracle import oracle  # noqa: E402
from tests.golden import cases  # noqa: E402

OUT = Path(__file__).with_name("steps.json")
THREADS = 8
TILES = {1: (8, 8), 2: (4, 4), 3: (7, 1), 4: (2, 2), 5: (2, 1)}


def tile_of(side: int) -> tuple[int, int]:
    """The pixels one wave of the tile tracer holds (pixelTileOf, artful_amd/csrc/hip/trace_kernel.h)."""
    return TILES.get(side, (1, 1))


def lane_slots(steps: np.ndarray) -> int:
    """steps: (side^2, H, W). The sum over the tile tracer's waves of its longest ray times its rays."""
    side = int(round(steps.shape[0] ** .5))
    tw, th = tile_of(side)
    h, w = steps.shape[1:]
    total = 0
    for y in range(0, h, th):
        block = steps[:, y:y + th].astype(np.int64)
        for x in range(0, w, tw):
            b = block[:, :, x:x + tw]
            total += int(b.max()) * b.size
    return total


def lane_slots_fast(steps: np.ndarray) -> int:
    """lane_slots, vectorised over the tiles (the frame padded with zero-step rays, which count no lanes)."""
    side = int(round(steps.shape[0] ** .5))
    tw, th = tile_of(side)
    n, h, w = steps.shape
    hp, wp = -(-h // th) * th, -(-w // tw) * tw
    padded = np.zeros((n, hp, wp), np.int64)
    padded[:, :h, :w] = steps
    valid = np.zeros((n, hp, wp), np.int64)
    valid[:, :h, :w] = 1
    tiles = padded.reshape(n, hp // th, th, wp // tw, tw)
    most = tiles.max(axis=(0, 2, 4))
    lanes = valid.reshape(n, hp // th, th, wp // tw, tw).sum(axis=(0, 2, 4))
    return int((most * lanes).sum())


def record(steps: np.ndarray, tile: bool = True) -> dict:
    counts = np.bincount(steps.ravel())
    at = np.nonzero(counts)[0]
    out = {"rays": int(steps.size), "steps": int(steps.astype(np.int64).sum()),
           "histogram": [[int(k), int(counts[k])] for k in at]}
    if tile:
        out["lane_slots"] = lane_slots_fast(steps)
    return out


def trace_steps(desc, cam, width, height, lw=None) -> np.ndarray:
    s = native.Scene(desc, cam)
    f, _ = s.frame(width, height)
    _, steps = oracle.trace(f, s.map(), s.fluid()[0], 32, lw=lw, steps=True, threads=THREADS)
    s.close()
    return steps


def config1() -> dict:
    desc, cam = cases.make(cases.CASES["config1"])
    return record(trace_steps(desc, cam, 320, 180))


def config2() -> dict:
    import bench  # the workload's view, as bench.py draws it
    desc, cam = bench.workload_scene(native)
    return record(trace_steps(desc, cam, bench.WIDTH, bench.HEIGHT))


def config3() -> dict:
    out = {}
    for spin in (0.0, .5, .9, .998):
        for incl in (17, 45, 75, 89):
            desc, cam, w, h = view_of(spin, incl, 50.0, 20.0, 3840, 2160, fov_y=.35)
            print(f"config3 spin {spin} inclination {incl}", flush=True)
            out[f"{spin},{incl}"] = record(trace_steps(desc, cam, w, h))
    return out


def config4() -> dict:
    views, _ = uneven_views(64, 4)
    per_view, hist = [], np.zeros(5001, np.int64)
    for k, (desc, cam, w, h) in enumerate(views):
        steps = trace_steps(desc, cam, w, h, lw=oracle.law(max_steps=5000))
        per_view.append(int(steps.astype(np.int64).sum()))
        hist += np.bincount(steps.ravel(), minlength=5001)[:5001]
        if k % 8 == 7:
            print(f"config4 view {k + 1} of 64", flush=True)
    at = np.nonzero(hist)[0]
    return {"rays": int(hist.sum()), "steps": int(sum(per_view)), "views": per_view,
            "histogram": [[int(k), int(hist[k])] for k in at]}


def config5() -> dict:
    from tests import kerr_autograd as K
    desc, cam, q = K.scene(K.HIDDEN)
    lw = oracle.law(max_steps=2000, escape_radius=100.0, disc="analytic", emissivity=q)
    return record(trace_steps(desc, cam, 512, 512, lw=lw))


MAKERS = {"config1": config1, "config2": config2, "config3": config3, "config4": config4, "config5": config5}


def main(names: list[str]) -> None:
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in names or list(MAKERS):
        print(name, flush=True)
        data[name] = MAKERS[name]()
    OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1:])
