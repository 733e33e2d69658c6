"""The batched, persistent tracer (include/artful/batch.h) against the HIP renderer, view by view.

Each view's HDR must equal what the renderer draws for a scene with the same description and camera:
bit for bit in FP32 (the same arithmetic per ray, whichever lane or wave traces it), under every
schedule, on every replay of the batch's graph. The renderer's FP32 HDR is itself held to the
reference's bit for bit (test_gpu_parity.py). FP64, mixed and FP32-with-FMA are held to the renderer
within 1e-6 relative on all but a few pixels (the two kernels may contract multiply-adds
differently)."""
from __future__ import annotations

import numpy as np
import pytest

from artful_amd import abi
from artful_amd.workloads import uneven_views, view_of

pytestmark = pytest.mark.gpu

# uneven sizes, 1 to 3 rays a side, sizes that are not multiples of the 8x8 tile, one pixel alone
SMALL = [view_of(.9, 80, 50, 20, 64, 36), view_of(0.0, 30, 30, 15, 37, 29, samples=2), view_of(.998, 60, 60, 30, 100, 21),
         view_of(.5, 85, 20, 10, 8, 8), view_of(.3, 5, 100, 40, 1, 1), view_of(.7, 45, 40, 25, 71, 66, samples=3)]


@pytest.fixture(scope="module")
def renderer(gpu):
    r = gpu.Renderer(64, 64)
    yield r
    r.close()


_drawn: dict = {}


def rendered(gpu, renderer, view, **tracing):
    """The renderer's HDR for one view (cached per view and tracing)."""
    desc, cam, w, h = view
    key = (bytes(desc), bytes(cam), w, h, tuple(sorted(tracing.items())))
    if key not in _drawn:
        s = gpu.Scene(desc, cam)
        renderer.resize(w, h)
        renderer.set_tracing(**tracing)
        assert renderer.draw(s)
        _drawn[key] = renderer.read_hdr()
        s.close()
    return _drawn[key]


def batch_images(gpu, views, schedule=None, **tracing):
    b = gpu.HoleBatch(views, gpu.tracing(**tracing))
    if schedule:
        b.set_schedule(*schedule)
    b.draw()
    images = [b.read_hdr(v) for v in range(len(views))]
    stats = b.stats()
    b.close()
    return images, stats


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fp32_painted_equals_renderer(gpu, renderer):
    tracing = dict(precision="fp32", disc="painted")
    images, stats = batch_images(gpu, SMALL, **tracing)
    for k, view in enumerate(SMALL):
        print("view", k, view[2:], flush=True)
        assert same_bits(images[k], rendered(gpu, renderer, view, **tracing)), f"view {k}"
    assert stats.rays == sum(w * h * d.samples ** 2 for d, _, w, h in SMALL)
    assert 0 < stats.steps <= stats.laneSlots


def test_fp32_tiny_spin_view_equals_oracle(gpu, oracle):
    """A batch with one view whose spin is below 2^-50 in size takes every view's drag by IEEE division
    (ArtfHipBatchArgs.dragInRange 0), the others by the in-range steps: the oracle's bits either way."""
    views = [view_of(1e-20, 60, 40, 20, 48, 27), view_of(.9, 70, 50, 20, 40, 24)]
    images, _ = batch_images(gpu, views, precision="fp32", disc="painted")
    for k, (desc, cam, w, h) in enumerate(views):
        s = gpu.Scene(desc, cam)
        f, _ = s.frame(w, h)
        want = oracle.trace(f, s.map(), s.fluid()[0], 32)
        assert same_bits(images[k], want), f"view {k}"
        s.close()


@pytest.mark.parametrize("precision", ["fp32_fast", "mixed", "fp64"])
def test_other_precisions_match_renderer(gpu, renderer, precision):
    tracing = dict(precision=precision, disc="painted")
    images, _ = batch_images(gpu, SMALL, **tracing)
    for k, view in enumerate(SMALL):
        want = rendered(gpu, renderer, view, **tracing).astype(np.float64)
        got = images[k].astype(np.float64)
        rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
        print(precision, "view", k, "bit-equal share", np.mean(got == want), "max rel", rel.max())
        assert (rel > 1e-6).mean() <= .005


def test_fp32_analytic_long_rays_equals_renderer(gpu, renderer):
    """The analytic disc, rays of up to 5000 steps, escape at r > 150."""
    tracing = dict(precision="fp32", disc="analytic", emissivity=2.5, max_steps=5000, escape_radius=150.0)
    images, stats = batch_images(gpu, SMALL, **tracing)
    for k, view in enumerate(SMALL):
        assert same_bits(images[k], rendered(gpu, renderer, view, **tracing)), f"view {k}"
    print("steps per ray", stats.steps / stats.rays, "active-lane fraction", stats.steps / stats.laneSlots)


@pytest.mark.parametrize("schedule", [(1, 1, 64, 1), (3, 5, 1, 3), (700, 64, 17, 16), (1024, 2, 8, 2), (1024, 2, 64, 8, 100), (5, 3, 20, 2, 7),
                                      (0, 0, 0, 0), (abi.HOLE_WORKGROUP_A_TILE, 1, 1, 1), (abi.HOLE_WORKGROUP_A_TILE, 2, 64, 8, 7)])
def test_schedules_give_the_same_image(gpu, renderer, schedule):
    """The queued kernel: one workgroup alone, odd chunks and refill thresholds, more workgroups than rays,
    the batch in slices of 100 and of 7 tiles (one launch each, queues cleared between); the kernel with a
    workgroup a tile, the default, whole and in slices of 7."""
    tracing = dict(precision="fp32", disc="painted")
    images, stats = batch_images(gpu, SMALL, schedule=schedule, **tracing)
    for k, view in enumerate(SMALL):
        assert same_bits(images[k], rendered(gpu, renderer, view, **tracing)), f"view {k} under {schedule}"
    assert stats.schedule.workgroups == (schedule[0] or abi.HOLE_WORKGROUP_A_TILE)
    assert stats.steps > 0 and stats.steps <= stats.laneSlots


def test_graph_replays(gpu, renderer):
    """Replays of the captured graph, with other device work between them, each give the same images,
    counts and a latency per view within the trace's span."""
    b = gpu.HoleBatch(SMALL, gpu.tracing(precision="fp32", disc="painted"))
    first, counts = None, None
    for replay in range(4):
        b.draw()
        images = [b.read_hdr(v) for v in range(len(SMALL))]
        s = b.stats()
        lat = b.latency()
        if first is None:
            first, counts = images, (s.rays, s.steps)
        else:
            assert all(same_bits(a, c) for a, c in zip(images, first)), f"replay {replay}"
            assert (s.rays, s.steps) == counts
        assert np.all(lat > 0) and np.all(lat <= s.traceMs + 1e-9) and s.traceMs <= s.batchMs * 1.01 + .05
        rendered(gpu, renderer, SMALL[replay % len(SMALL)], precision="fp32", disc="analytic")  # other work between replays
    b.close()


@pytest.mark.parametrize("workgroups", [abi.HOLE_WORKGROUP_A_TILE, 1024])
def test_graph_replays_in_slices(gpu, renderer, workgroups):
    """The batch in slices of 7 tiles, which take turns on two streams (with a set of queues each, for the
    queued kernel): every replay of its graph, with other device work between them, gives the renderer's
    images and the same counts."""
    tracing = dict(precision="fp32", disc="painted")
    b = gpu.HoleBatch(SMALL, gpu.tracing(**tracing))
    b.set_schedule(workgroups, 2, 64, 8, 7)
    counts = None
    for replay in range(3):
        b.draw()
        s = b.stats()
        if counts is None:
            counts = (s.rays, s.steps, s.laneSlots)
        assert (s.rays, s.steps, s.laneSlots) == counts, f"replay {replay}"
        for k, view in enumerate(SMALL):
            assert same_bits(b.read_hdr(k), rendered(gpu, renderer, view, **tracing)), f"view {k}, replay {replay}"
        rendered(gpu, renderer, SMALL[replay % len(SMALL)], precision="fp32", disc="analytic")  # other work between replays
    b.close()


def test_tune_keeps_the_image(gpu, renderer):
    tracing = dict(precision="fp32", disc="painted")
    b = gpu.HoleBatch(SMALL, gpu.tracing(**tracing))
    best = b.tune(rounds=1)
    assert (best.workgroups >= 1 or best.workgroups == abi.HOLE_WORKGROUP_A_TILE) and 1 <= best.chunkTiles <= 64 and 1 <= best.refill <= 64 and 1 <= best.stepsPerCheck <= 16
    b.draw()
    s = b.stats()
    assert (s.schedule.workgroups, s.schedule.chunkTiles, s.schedule.refill, s.schedule.stepsPerCheck) == \
        (best.workgroups, best.chunkTiles, best.refill, best.stepsPerCheck)
    for k, view in enumerate(SMALL):
        assert same_bits(b.read_hdr(k), rendered(gpu, renderer, view, **tracing)), f"view {k}"
    b.close()


def test_uneven_views_workload(gpu, renderer):
    """Eight views of the uneven-views workload at 128..256 pixels a side, painted disc, 5000 steps:
    two of them against the renderer, and the lanes kept busy."""
    views, _ = uneven_views(count=8, seed=4, sizes=(128, 256))
    tracing = dict(precision="fp32", disc="painted", max_steps=5000)
    images, stats = batch_images(gpu, views, **tracing)
    for k in (0, 5):
        assert same_bits(images[k], rendered(gpu, renderer, views[k], **tracing)), f"view {k}"
    busy = stats.steps / stats.laneSlots
    print("rays", stats.rays, "steps per ray", stats.steps / stats.rays, "active-lane fraction", busy, "trace ms", stats.traceMs)
    assert busy > .7


def test_c_abi_rejects_bad_schedules(gpu):
    L = gpu.lib()
    a = abi.HipBatchArgs()
    assert L.artfHipTraceViews(a, None) != 0  # no views


def test_huge_frame_in_slices(gpu, renderer, oracle):
    """A frame of 8x8 rays a pixel (640 x 360: 14.7 M rays) as a batch of one view in slices of 20,000
    tiles (12 launches), against the renderer's one launch, and bands of its rows against the oracle."""
    view = view_of(.998, 60, 60, 30, 640, 360, fov_y=.35, samples=8)
    tracing = dict(precision="fp32", disc="painted")
    images, stats = batch_images(gpu, [view], schedule=(0, 2, 64, 8, 20000), **tracing)
    assert stats.rays == 640 * 360 * 64
    assert same_bits(images[0], rendered(gpu, renderer, view, **tracing))
    s = gpu.Scene(view[0], view[1])
    f, _ = s.frame(640, 360)
    for band in ((0, 2), (176, 184), (358, 360)):
        want = oracle.trace(f, s.map(), s.fluid()[0], 32, rows=band)
        assert same_bits(images[0][:, band[0]:band[1]], want[:, band[0]:band[1]]), band
    s.close()
