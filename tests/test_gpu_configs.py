"""The two workloads the other GPU tests only sample, at their real sizes.

The precision study: 3840x2160, one ray a pixel, spins {0, .5, .9, .998} x inclinations {17, 45, 75, 89},
eye at r 50, painted disc ISCO..20, Doppler 1. Each image's FP32 and mixed traces against its FP64 trace
(relative error, floor 1e-3): the mean and the share of pixels past 1e-3 held to about twice what was
measured for that image, the mixed tracer nearer FP64 than FP32 in both; and a band of rows of each image
against the oracle, the FP32 trace bit for bit and the FP64 one within 1e-5.

The batch: the 64 uneven views (seed 4: 128..1024 pixels a side, spin U[0, .998], inclination U[5, 85],
eye r U[20, 100], disc outer U[10, 40]), painted disc, rays of up to 5000 steps, in one persistent launch.
Every view finite and non-negative, the counts consistent; the twelve smallest views whole, and a band of
rows of each of the six largest, bit for bit equal to the oracle tracing that view on a host scene of the
same description (its own ring tables and fluid, made by the host code the reference's digests pin, where
the batch made its own on the device).

The huge frame: 3840x2160 at 8x8 rays a pixel (530,841,600 rays), through the renderer's one launch of
the tile tracer (which sums each pixel's 64 rays in its wave) and through a batch of one view in
watchdog-sized launches of 2^20 tiles: the two equal bit for bit, and bands of rows equal to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from artful_amd.workloads import uneven_views, view_of

pytestmark = pytest.mark.gpu

W4K, H4K = 3840, 2160
# measured on MI355X (mean relative error, share of pixels past 1e-3) per (spin, inclination): FP32, mixed
MEASURED = {
    (0.0, 17): (1.50e-4, .0080, 5.65e-5, .0027), (0.0, 45): (1.24e-4, .0042, 4.83e-5, .0022),
    (0.0, 75): (8.43e-5, .0037, 4.18e-5, .0021), (0.0, 89): (1.66e-3, .0047, 5.21e-5, .0030),
    (.5, 17): (1.04e-4, .0048, 3.62e-5, .0014), (.5, 45): (1.47e-4, .0033, 1.18e-4, .0016),
    (.5, 75): (8.15e-5, .0037, 4.04e-5, .0021), (.5, 89): (4.40e-4, .0047, 4.90e-5, .0030),
    (.9, 17): (6.93e-5, .0026, 2.62e-5, .0008), (.9, 45): (4.41e-5, .0018, 2.68e-5, .0009),
    (.9, 75): (7.56e-5, .0037, 3.22e-5, .0013), (.9, 89): (5.69e-5, .0042, 2.84e-5, .0020),
    (.998, 17): (3.69e-5, .0023, 2.68e-5, .0008), (.998, 45): (2.94e-5, .0012, 2.16e-5, .0008),
    (.998, 75): (3.60e-5, .0022, 2.37e-5, .0009), (.998, 89): (3.51e-5, .0022, 2.53e-5, .0014),
}


def relative(a, ref):
    return np.abs(a.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-3)


@pytest.fixture(scope="module")
def renderer4k(gpu):
    r = gpu.Renderer(W4K, H4K)
    yield r
    r.close()


@pytest.mark.parametrize("spin", [0.0, .5, .9, .998])
def test_precision_study_4k(spin, gpu, renderer4k, oracle):
    s = gpu.Scene()
    band = (H4K // 2 - 4, H4K // 2 + 4)
    for incl in (17, 45, 75, 89):
        desc, cam, _, _ = view_of(spin, incl, 50.0, 20.0, W4K, H4K, fov_y=.35)
        s.set_hole(desc)
        s.set_camera(cam)
        hdr = {}
        for p in ("fp64", "fp32", "mixed"):
            renderer4k.set_tracing(p)
            assert renderer4k.draw(s)
            hdr[p] = renderer4k.read_hdr()
        ref = hdr["fp64"].astype(np.float64)
        err = {}
        for p in ("fp32", "mixed"):
            rel = relative(hdr[p], ref)
            err[p] = (float(rel.mean()), float((rel > 1e-3).mean()), float(np.quantile(rel, .999)), float(rel.max()))
            assert np.isfinite(hdr[p]).all() and (hdr[p] >= 0).all()
        print(f"4K spin {spin} inclination {incl}: fp32 mean/share/p99.9/max {err['fp32']}, mixed {err['mixed']}", flush=True)
        f32_mean, f32_share, mix_mean, mix_share = MEASURED[(spin, incl)]
        assert err["fp32"][0] <= 2 * f32_mean and err["fp32"][1] <= 2 * f32_share
        assert err["mixed"][0] <= 2 * mix_mean and err["mixed"][1] <= 2 * mix_share
        assert err["mixed"][0] < err["fp32"][0] and err["mixed"][1] < err["fp32"][1]
        # a band of rows against the oracle: FP32 bit for bit, FP64 within 1e-5 (the GPU fuses multiply-adds)
        f, _ = s.frame(W4K, H4K)
        m, fl = s.map(), s.fluid()[0]
        want32 = oracle.trace(f, m, fl, 32, rows=band)[:, band[0]:band[1]]
        np.testing.assert_array_equal(hdr["fp32"][:, band[0]:band[1]].view(np.uint32), want32.view(np.uint32))
        want64 = oracle.trace(f, m, fl, 64, s.frame64(W4K, H4K), rows=band)[:, band[0]:band[1]].astype(np.float64)
        rel64 = relative(hdr["fp64"][:, band[0]:band[1]], want64)
        assert np.median(rel64) < 1e-6 and (rel64 > 1e-5).mean() < 2e-3, (np.median(rel64), (rel64 > 1e-5).mean())
    s.close()


def test_batch_of_64_uneven_views_against_oracle(gpu, oracle):
    views, params = uneven_views(64, 4)
    b = gpu.HoleBatch(views, gpu.tracing(precision="fp32", disc="painted", max_steps=5000))
    b.draw()  # the default schedule: a workgroup a tile
    st = b.stats()
    images = [b.read_hdr(v) for v in range(len(views))]
    b.set_schedule(1024, 1, 16, 4)  # the queued kernel, idle lanes refilled 16 at a time
    b.draw()
    queued = b.stats()
    for k in range(len(views)):
        np.testing.assert_array_equal(b.read_hdr(k).view(np.uint32), images[k].view(np.uint32), err_msg=f"view {k}, queued")
    b.close()
    assert queued.steps == st.steps
    rays = sum(w * h * d.samples ** 2 for d, _, w, h in views)
    assert st.views == 64 and st.rays == rays and 0 < st.steps <= st.laneSlots
    assert 10 < st.steps / st.rays < 60, st.steps / st.rays
    for k, img in enumerate(images):
        assert img.shape == (3, views[k][3], views[k][2])
        assert np.isfinite(img).all() and (img >= 0).all() and img.max() > 0, f"view {k}"
    by_size = sorted(range(64), key=lambda k: views[k][2] * views[k][3])
    s = gpu.Scene()
    for k in by_size[:12] + by_size[-6:]:
        desc, cam, w, h = views[k]
        s.set_hole(desc)  # the same description: its rings and fluid made again on the host
        s.set_camera(cam)
        f, _ = s.frame(w, h)
        whole = k in by_size[:12]
        rows = (0, h) if whole else (h // 2 - 8, h // 2 + 8)
        print(f"batch view {k}: {w}x{h} spin {params[k]['spin']:.3f} {'whole' if whole else f'rows {rows}'}", flush=True)
        want = oracle.trace(f, s.map(), s.fluid()[0], 32, lw=oracle.law(max_steps=5000), rows=rows)
        np.testing.assert_array_equal(images[k][:, rows[0]:rows[1]].view(np.uint32), want[:, rows[0]:rows[1]].view(np.uint32))
    s.close()


def test_huge_frame_4k_64_rays(gpu, oracle):
    import bench  # the workload's view, as bench.py draws it
    desc, cam = bench.workload_scene(gpu, samples=8)
    s = gpu.Scene(desc, cam)
    r = gpu.Renderer(W4K, H4K)
    r.set_tracing("fp32")
    assert r.draw(s)
    whole = r.read_hdr()
    st = r.hole_stats()
    r.close()
    assert st.rays == W4K * H4K * 64 and 0 < st.steps <= st.laneSlots
    assert np.isfinite(whole).all() and (whole >= 0).all()
    b = gpu.HoleBatch([(desc, cam, W4K, H4K)], gpu.tracing(precision="fp32", disc="painted"))
    b.set_schedule(0, 2, 64, 8, 1 << 20)
    b.draw()
    sliced = b.read_hdr(0)
    bs = b.stats()
    b.close()
    assert bs.rays == st.rays and bs.steps == st.steps
    np.testing.assert_array_equal(sliced.view(np.uint32), whole.view(np.uint32))
    f, _ = s.frame(W4K, H4K)
    for band in ((0, 1), (H4K // 2 - 1, H4K // 2 + 1), (H4K - 1, H4K)):
        print(f"huge frame rows {band}", flush=True)
        want = oracle.trace(f, s.map(), s.fluid()[0], 32, rows=band)
        np.testing.assert_array_equal(whole[:, band[0]:band[1]].view(np.uint32), want[:, band[0]:band[1]].view(np.uint32))
    s.close()
