"""Kerr black-hole tracer throughput on one MI355X, with its roofline and the reference CPU path beside it.

Workload: 1920x1080, 2x2 stratified rays a pixel (8,294,400 rays a frame), spin .998, observer r=60M,
inclination 60 deg, painted disc from the last stable orbit to 30M, Doppler 1; HDR, bloom at the
reference's levels, tone map. A step is one frame through artfRendererDraw (trace, resolve, bloom, tone
map) with the map and fluid already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line. value = rays per frame x K / (wall time of the K frames), FP32 tracer.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WIDTH, HEIGHT, SAMPLES = 1920, 1080, 2
SPIN, RADIUS, INCLINATION, OUTER = .998, 60.0, 60.0, 30.0
FLOPS_PER_STEP = {True: 147, False: 122}  # per Nystrom step, spin != 0 / spin == 0 (+ - x / one each, an FMA two)
PEAK_FP32_TFLOPS = 157.3                    # MI355X vector FP32, published
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_scene(native, samples=SAMPLES):
    i = math.radians(INCLINATION)
    desc = native.hole_desc(spin=SPIN, axis=(0, 1, 0), discInner=0.0, discOuter=OUTER, doppler=1.0, samples=samples)
    up = (0, 1, 0) if math.pi / 2 - i <= 1 else (0, 0, -1)
    cam = native.camera((0, RADIUS * math.cos(i), RADIUS * math.sin(i)), (0, 0, 0), up, .35, .1)
    return desc, cam


def settle_frames(renderer, scene, seconds):
    """Frames drawn untimed for `seconds` of wall time: the frame path's first few dozen frames after an
    idle spell run ~10% slower on the GPU's own clock (measured: 0.98 against 0.885 ms a frame at 1080p,
    the same after 2 s of idle), so the timed frames follow a settled GPU. Returns how many were drawn."""
    n, t_end = 0, time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        if not renderer.draw(scene):
            raise RuntimeError("draw failed")
        n += 1
    renderer.hole_stats()
    return n


def time_frames(native, renderer, scene, precision, steps, warmup, settle_s=0.0):
    """K frames back to back (artfRendererDraw queues a frame; the last one is waited for by the stats
    read after the loop, inside the timed region), after `settle_s` seconds of untimed frames and then W
    untimed warm-up frames, their steps not counted (ArtfHoleTracing.uncounted: the trace skips the
    counting's wave reductions and atomics, which are the stats', not the image's); then frames one at a
    time, counted, for each stage's GPU time and the step counts (the same rays: the same steps)."""
    renderer.set_tracing(precision, count_steps=False)
    settled = settle_frames(renderer, scene, settle_s) if settle_s > 0 else 0
    for _ in range(warmup):
        if not renderer.draw(scene):
            raise RuntimeError("draw failed")
    renderer.hole_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        if not renderer.draw(scene):
            raise RuntimeError("draw failed")
    st = renderer.hole_stats()  # waits for the last frame
    wall = time.perf_counter() - t0
    # the same frames on the GPU's own clock: the first one's trace start to the last one's tone-map end
    assert st.frames == steps, (st.frames, steps)
    gpu_span = st.spanMs
    renderer.set_tracing(precision)
    trace, post, slots, nsteps = [], [], 0, 0
    staged = max(3, min(steps, 10))
    for _ in range(staged):
        if not renderer.draw(scene):
            raise RuntimeError("draw failed")
        st = renderer.hole_stats()
        trace.append(st.traceMs)
        post.append(st.resolveMs + st.bloomMs + st.toneMs)
        nsteps += st.steps
        slots += st.laneSlots
    return {"wall_s": wall, "gpu_span_ms": gpu_span, "settle": {"seconds": settle_s, "frames": settled}, "trace_ms": float(np.mean(trace)), "post_ms": float(np.mean(post)), "rays": int(st.rays),
            "steps_per_frame": nsteps / staged, "active_lanes": nsteps / max(slots, 1), "bloom_levels": st.bloomLevels}


def kernel_ms(native, scene, precision, launches=20, warmup=100):
    """The trace kernel alone on the workload's frame: artfHipTrace launched back to back on a stream of
    its own, bracketed by two HIP events on that stream; milliseconds a launch. The kernel writes the
    HDR planes (each pixel's rays summed in its lanes) and does not count its steps, as in the timed
    frames. The untimed launches first (about 0.1 s) bring the GPU's clock up from where an idle spell
    left it."""
    import ctypes as C
    from artful_amd import abi
    L = native.lib()
    f, _ = scene.frame(WIDTH, HEIGHT)
    a = abi.HipTraceArgs()
    a.frame = f
    a.frame64 = scene.frame64(WIDTH, HEIGHT)
    a.precision = abi.PRECISION[precision]
    host = [np.ascontiguousarray(scene.map()), np.ascontiguousarray(scene.fluid()[0])]
    held = [L.artfHipMalloc(h.nbytes) for h in host] + [L.artfHipMalloc(WIDTH * HEIGHT * 4) for _ in range(3)]
    stream, start, stop = L.artfHipStreamCreate(), L.artfHipEventCreate(), L.artfHipEventCreate()
    try:
        if not all(held) or not (stream and start and stop):
            raise RuntimeError("no room for the kernel timing")
        for d, h in zip(held, host):
            if L.artfHipUpload(d, h.ctypes.data, h.nbytes, stream):
                raise RuntimeError("upload failed")
        a.map, a.fluid = held[0], held[1]
        for c in range(3):
            a.out[c] = held[2 + c]
        a.counters = None
        for _ in range(warmup):
            if L.artfHipTrace(C.byref(a), stream):
                raise RuntimeError("trace failed")
        if L.artfHipStreamSync(stream) or L.artfHipEventRecord(start, stream):
            raise RuntimeError("event failed")
        for _ in range(launches):
            if L.artfHipTrace(C.byref(a), stream):
                raise RuntimeError("trace failed")
        if L.artfHipEventRecord(stop, stream) or L.artfHipStreamSync(stream):
            raise RuntimeError("trace failed")
        return L.artfHipEventMs(start, stop) / launches
    finally:
        for d in held:
            if d:
                L.artfHipFree(d)
        L.artfHipEventDestroy(start), L.artfHipEventDestroy(stop), L.artfHipStreamDestroy(stream)


def rel_error(a, ref, floor=1e-3):
    rel = np.abs(a.astype(np.float64) - ref) / np.maximum(np.abs(ref), floor)
    return {"max": float(rel.max()), "mean": float(rel.mean()), "share_over_1e-3": float((rel > 1e-3).mean())}


def traffic_from_profiles():
    """HBM bytes per trace launch from the committed rocprofv3 --pmc summary, or None."""
    p = ROOT / "profiles" / "trace_1080p_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"bytes per launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(native, frames=12, warmup=2):
    """The reference's own software renderer (oracle/_ref, built from its sources) on the workload's
    frame itself: 1920x1080 at 2x2 rays a pixel, the same view, min(CPUs, 16) threads (the reference's
    rule, src/core/threads.c:54, 144), AVX2; the median over `frames` frames after `warmup`, of its
    lensMs (trace + bloom, software.c:673-675) and of its rasterMs (the tiles' pass, whose work with no
    triangles is the tone map, kernels.inc:1616-1617). It traces the very rays the GPU does (its HDR is
    the GPU's bit for bit), so the steps a ray are the GPU frame's. Where the reference library is absent,
    the oracle (this project's C restatement, bit-equal to it) is timed on a band of rows of the same
    frame instead."""
    from oracle import oracle as O
    desc, cam = workload_scene(native)
    cores = os.cpu_count() or 1
    threads = min(len(os.sched_getaffinity(0)), 16)
    rays = WIDTH * HEIGHT * SAMPLES * SAMPLES
    if O.reference_available():
        import ctypes as C
        L = O.reference()
        t0 = time.perf_counter()
        scene = O.RefScene(desc, cam)  # the reference's map paint and fluid warm-up: once per scene, reported apart
        setup_s = time.perf_counter() - t0
        r = L.refRendererNew(WIDTH, HEIGHT, 0, threads)
        lens, raster, cpu = C.c_double(), C.c_double(), C.c_double()
        got = []
        for k in range(warmup + frames):
            L.refRendererDraw(r, scene.ptr)
            L.refRendererStats(r, C.byref(lens), C.byref(raster), C.byref(cpu))
            if k >= warmup:
                got.append((lens.value, raster.value, cpu.value))
        description = L.refRendererDescription(r).decode()
        L.refRendererFree(r)
        scene.close()
        lens_ms, raster_ms, cpu_ms = (float(np.median([g[i] for g in got])) for i in range(3))
        return {"value": rays / (lens_ms / 1e3), "unit": "rays/s", "cores": threads, "kind": "reference",
                "sample": f"reference software renderer ({description}; {threads} threads of {cores} CPUs): the whole "
                          f"{WIDTH}x{HEIGHT} frame of the same view at 2x2 rays/pixel, trace+bloom (lensMs), median of {frames} "
                          f"frames after {warmup}",
                "lens_ms": lens_ms, "tone_ms": raster_ms, "frame_ms": cpu_ms, "scene_setup_s": setup_s,
                "rays_per_s_with_tone_map": rays / ((lens_ms + raster_ms) / 1e3)}
    scene = native.Scene(desc, cam)
    frame, _ = scene.frame(WIDTH, HEIGHT)
    m, (fl, _, _) = scene.map(), scene.fluid()
    band = (HEIGHT // 2 - 32, HEIGHT // 2 + 32)
    t0 = time.perf_counter()
    O.trace(frame, m, fl, 32, rows=band, threads=threads)
    dt = time.perf_counter() - t0
    band_rays = WIDTH * (band[1] - band[0]) * SAMPLES * SAMPLES
    return {"value": band_rays / dt, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"oracle restatement, {threads} threads: rows {band[0]}..{band[1]} of the same frame, 2x2 rays/pixel, trace only"}


def adjoint_fit(native, iterations=200):
    """The adjoint workload: 512x512, one ray a pixel, the analytic disc; the adjoint's cost against the
    forward trace, and a 200-step Adam fit of (spin, inclination, disc inner, outer, q) from the start
    point to the image of the hidden parameters."""
    from tests import kerr_autograd as K
    ctx = native.HoleAdjoint(512, 512)
    target = K.render(ctx, K.HIDDEN)
    fwd, adj = [], []
    for _ in range(5):
        img = K.render(ctx, K.START)
        fwd.append(ctx.times()[0])
        K.gradient(ctx, K.START, (2 * (img - target) / img.size).astype(np.float32))
        adj.append(ctx.times()[1])
    t0 = time.perf_counter()
    history = K.fit(ctx, target, iterations=iterations)
    wall = time.perf_counter() - t0
    ctx.close()
    least, params = K.best(history)
    return {"forward_ms": float(np.median(fwd)), "adjoint_ms": float(np.median(adj)),
            "adjoint_over_forward": float(np.median(adj) / np.median(fwd)),
            "fit": {"iterations": iterations, "wall_s": wall, "loss_first": history[0][0], "loss_last": history[-1][0],
                    "loss_best": least, "params_best": [float(v) for v in params], "params_last": [float(v) for v in history[-1][1]],
                    "hidden": list(K.HIDDEN), "start": list(K.START),
                    "params_error": {n: float(v - h) for n, v, h in zip(K.PARAMS, params, K.HIDDEN)},
                    "params_error_last": {n: float(v - h) for n, v, h in zip(K.PARAMS, history[-1][1], K.HIDDEN)},
                    "result": "the iterate of least loss",
                    "preconditioner": "Gauss-Newton whitening from forward central differences, refreshed every 25 iterations; iterates held within bounds"}}


def per_view_tiles(native, views, trc, reps=3):
    """The same views traced by the tile kernel (artfHipTrace), one launch per view back to back on one
    stream: GPU ms of the whole sequence (best of reps), its steps and lane-slots. Analytic disc only
    (its tables are the map's shared colours and sky, so one map serves every view)."""
    import ctypes as C
    from artful_amd import abi
    L = native.lib()
    themap = native.Scene(*views[0][:2]).map()
    dmap = L.artfHipMalloc(themap.nbytes)
    stream = L.artfHipStreamCreate()
    L.artfHipUpload(dmap, themap.ctypes.data, themap.nbytes, stream)
    counters = L.artfHipMalloc(2 * 1024 * 8)
    args, planes = [], []
    for desc, cam, w, h in views:
        f, f64 = native.frames_for(desc, cam, w, h)
        a = abi.HipTraceArgs()
        a.frame, a.frame64, a.map, a.fluid, a.counters = f, f64, dmap, dmap, counters
        for c in range(3):
            buf = L.artfHipMalloc(w * h * 4)
            planes.append(buf)
            a.out[c] = buf
        a.precision, a.maxSteps, a.escapeRadius, a.disc, a.emissivity = trc.precision, trc.maxSteps, trc.escapeRadius, trc.disc, trc.emissivity
        args.append(a)
    e0, e1 = L.artfHipEventCreate(), L.artfHipEventCreate()
    best, steps, slots = 1e30, 0, 0
    host = np.zeros(2 * 1024, np.uint64)
    for _ in range(reps):
        L.artfHipMemset(counters, 0, 2 * 1024 * 8, stream)
        L.artfHipEventRecord(e0, stream)
        for a in args:
            if L.artfHipTrace(C.byref(a), stream):
                raise RuntimeError("artfHipTrace failed")
        L.artfHipEventRecord(e1, stream)
        L.artfHipStreamSync(stream)
        best = min(best, L.artfHipEventMs(e0, e1))
        L.artfHipDownload(host.ctypes.data, counters, host.nbytes, stream)
        steps, slots = int(host[:1024].sum()), int(host[1024:].sum())
    for b in planes + [dmap, counters]:
        L.artfHipFree(b)
    L.artfHipEventDestroy(e0), L.artfHipEventDestroy(e1), L.artfHipStreamDestroy(stream)
    return {"ms": best, "steps": steps, "lane_slots": slots}


def batched_views(native, count=64, seed=4, max_steps=5000, draws=5, schedule=None):
    """The batch workload: 64 views (seed 4) of 128..1024 pixels a side, spin U[0,.998], inclination
    U[5,85], eye r U[20,100], disc outer U[10,40], each with its own rings and fluid; rays of up to 5000
    steps; one launch replayed from a HIP graph, schedule from the tuner (a workgroup a tile, or the
    persistent queued kernel with its workgroups, chunk, refill and check interval), its steps not counted
    (a counted batch of the same views gives the step counts). Beside it, the
    analytic disc through the same batch and through the tile kernel one launch per view."""
    from artful_amd.workloads import uneven_views
    views, _ = uneven_views(count, seed)
    out = {"views": count, "seed": seed, "max_steps": max_steps}
    for disc in ("painted", "analytic"):
        trc = native.tracing(precision="fp32", disc=disc, max_steps=max_steps)
        # the timed batch does not count its steps (as the timed frames); a counted one gives the counts
        counted = native.HoleBatch(views, trc)
        counted.draw()
        cst = counted.stats()
        counted.close()
        t0 = time.perf_counter()
        b = native.HoleBatch(views, native.tracing(precision="fp32", disc=disc, max_steps=max_steps, count_steps=False))
        setup = time.perf_counter() - t0

        def run(n):
            ms, trace = [], []
            for _ in range(n):
                b.draw()
                st = b.stats()
                ms.append(st.batchMs)
                trace.append(st.traceMs)
            return st, float(np.median(ms)), float(np.median(trace))

        b.draw()
        _, default_ms, _ = run(draws)
        t0 = time.perf_counter()
        if schedule:
            b.set_schedule(*schedule)
            sched = b.stats().schedule
        else:
            sched = b.tune()
        tune_s = time.perf_counter() - t0
        st, batch_ms, trace_ms = run(draws)
        lat = b.latency()
        b.close()
        leg = {"setup_s": setup, "rays": int(st.rays), "steps_per_ray": cst.steps / cst.rays,
               "graph_ms": batch_ms, "trace_ms": trace_ms, "rays_per_s": st.rays / (batch_ms * 1e-3),
               "active_lane_fraction": cst.steps / max(cst.laneSlots, 1), "default_schedule_ms": default_ms,
               "schedule": {"workgroups": sched.workgroups, "chunk_tiles": sched.chunkTiles, "refill": sched.refill,
                            "steps_per_check": sched.stepsPerCheck}, "tune_s": tune_s,
               "latency_ms": {"min": float(lat.min()), "median": float(np.median(lat)), "max": float(lat.max()),
                              "per_view": [round(float(v), 4) for v in lat]}}
        if disc == "analytic":
            tiles = per_view_tiles(native, views, trc)
            leg["tile_kernel_per_view"] = {"ms": tiles["ms"], "rays_per_s": st.rays / (tiles["ms"] * 1e-3),
                                           "active_lane_fraction": tiles["steps"] / max(tiles["lane_slots"], 1),
                                           "launches": count}
            leg["speedup_over_tile_kernel"] = tiles["ms"] / batch_ms
        out[disc] = leg
        log(f"batch {disc}: {leg}")
    return out


def animation(native, frames=100, warmup=50):
    """The reference driver's loop (tools/blackhole.c: time advanced .2 a frame, one fluid step) on the
    1080p view: artfSceneSetBlackHole then artfRendererDraw, per frame, wall clock, with the disc's map
    and fluid on the host (the reference's way: the fluid stepped by the CPU, then uploaded) and on the
    device (artfSceneSetDiscOnDevice); also the scene's first set-up (map paint + fluid warm-up) each
    way, and whether the two give the same frames. Each way first draws `warmup` frames untimed (the GPU's
    clock comes up from idle over the first ~30), then times `frames`."""
    desc, cam = workload_scene(native)
    r = native.Renderer(WIDTH, HEIGHT)
    r.set_tracing("fp32")
    out = {"frames": frames, "warmup": warmup}
    last = {}
    for where in ("host", "device"):
        t0 = time.perf_counter()
        s = native.Scene(cam=cam)
        if where == "device":
            s.set_disc_on_device(0)
        desc.time = 0.0
        s.set_hole(desc)
        setup = time.perf_counter() - t0
        r.draw(s)
        hole_ms = 0.0
        for k in range(1, warmup + frames + 1):
            if k == warmup + 1:
                r.hole_stats()  # waits for the last warm-up frame
                t0 = time.perf_counter()
                hole_ms = 0.0
            desc.time = .2 * k
            t1 = time.perf_counter()
            s.set_hole(desc)
            hole_ms += (time.perf_counter() - t1) * 1e3
            if not r.draw(s):
                raise RuntimeError("draw failed")
        r.hole_stats()  # waits for the last frame
        wall = time.perf_counter() - t0
        last[where] = r.read_hdr()
        out[where] = {"setup_s": setup, "ms_per_frame": wall / frames * 1e3, "set_hole_ms_per_frame": hole_ms / frames,
                      "trace_ms": r.hole_stats().traceMs}
        s.close()
        log(f"animation {where}: {out[where]}")
    r.close()
    out["same_frames"] = bool(np.array_equal(last["host"].view(np.uint32), last["device"].view(np.uint32)))
    out["speedup"] = out["host"]["ms_per_frame"] / out["device"]["ms_per_frame"]
    return out


# Artful's published black-hole figures (its README.md:381-390, `artful-render --scene blackhole` at 1920x1080,
# GPU frames timed to the last pixel of a run): ms a frame on an AMD Vega 8 iGPU (Vulkan) and on the Ryzen 3
# 3200G's 4 cores (its software renderer)
PUBLISHED_BLACKHOLE_MS = {"vega8_vulkan": 23.0, "ryzen_3200g_4_cores_software": 141.0}


def render_scene(native, frames=300, warmup=50, width=1920, height=1080):
    """The scene of Artful's published black-hole figures: tools/render.c's black-hole scene with its
    defaults (render.c:398: the eye 74.1 masses away and 3.44 degrees above the disc, a 19 degree field of
    view, spin .6, one ray a pixel, no Doppler; render.c:335-360 places the eye), at 1920x1080, its time
    going on by .2 a frame so that each frame also takes one step of the disc's fluid (render.c:551-553),
    and the frames timed as a run to the last pixel with the read-back taken out (render.c:545-575): with
    the fluid stepped on the CPU (the reference's way, the same host code for every backend) and on the
    GPU (artfSceneSetDiscOnDevice), beside the published milliseconds a frame."""
    f32 = np.float32
    tilt = f32(3.44) * f32(3.14159265) / f32(180.0)
    dist = f32(74.1)
    eye = (0.0, float(np.sin(tilt, dtype=f32) * dist), float(np.cos(tilt, dtype=f32) * dist))
    fov = float(f32(19.0) * f32(3.14159265) / f32(180.0))
    cam = native.camera(eye, (0, 0, 0), (0, 1, 0), fov, .1)
    r = native.Renderer(width, height)
    r.set_tracing("fp32")
    out = {"width": width, "height": height, "rays": width * height, "frames": frames, "warmup": warmup,
           "published_ms": PUBLISHED_BLACKHOLE_MS}
    last = {}
    for where in ("host", "device"):
        s = native.Scene(cam=cam)
        if where == "device":
            s.set_disc_on_device(0)
        desc = native.defaults()
        desc.time, desc.doppler, desc.samples, desc.exposure, desc.spin, desc.shutter = 0.0, 0.0, 1, 1.0, .6, 0.0
        s.set_hole(desc)
        t = f32(0.0)
        for k in range(warmup + frames):
            if k == warmup:
                r.read_pixels()  # the warm-up frames done: the run starts from an idle queue
                t0 = time.perf_counter()
            t = t + f32(.2)
            desc.time = float(t)
            s.set_hole(desc)
            if not r.draw(s):
                raise RuntimeError("draw failed")
        r.read_pixels()
        wall = time.perf_counter() - t0
        t1 = time.perf_counter()
        last[where] = r.read_pixels()
        wall -= time.perf_counter() - t1
        st = r.hole_stats()
        out[where] = {"ms_per_frame": wall / frames * 1e3, "gpu_ms_last_frame": st.traceMs + st.resolveMs + st.bloomMs + st.toneMs}
        s.close()
        log(f"render scene {where}: {out[where]}")
    r.close()
    out["same_frames"] = bool(np.array_equal(last["host"], last["device"]))
    out["vs_published"] = {k: v / out["device"]["ms_per_frame"] for k, v in PUBLISHED_BLACKHOLE_MS.items()}
    return out


EXTENDED_LAW = dict(disc="analytic", emissivity=3.0, jitter=2, bloom_levels=5)


def extended_form(native, renderer=None, scene=None, steps=100, warmup=30):
    """The 1080p workload with its extended ray law: 2x2 rays a pixel at places jittered from seed 2
    inside their strata, the analytic disc (emissivity (inner/r)^3), 5 bloom levels. FP32 and mixed
    frames back to back, and each one's relative HDR error against the FP64 frame of the same law."""
    own = renderer is None
    if own:
        desc, cam = workload_scene(native)
        scene = native.Scene(desc, cam)
        renderer = native.Renderer(WIDTH, HEIGHT)
    out = {"law": dict(EXTENDED_LAW)}
    renderer.set_tracing("fp64", **EXTENDED_LAW)
    if not renderer.draw(scene):
        raise RuntimeError("draw failed")
    ref64 = renderer.read_hdr().astype(np.float64)
    for p in ("fp32", "mixed"):
        renderer.set_tracing(p, **EXTENDED_LAW)
        for _ in range(warmup):
            if not renderer.draw(scene):
                raise RuntimeError("draw failed")
        renderer.hole_stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            if not renderer.draw(scene):
                raise RuntimeError("draw failed")
        st = renderer.hole_stats()  # waits for the last frame
        wall = time.perf_counter() - t0
        out[p] = {"ms_per_frame": wall / steps * 1e3, "rays_per_s": st.rays * steps / wall, "bloom_levels": st.bloomLevels,
                  "error_vs_fp64": rel_error(renderer.read_hdr(), ref64)}
    # the same law with the sky's stars off: the disc and the dust of the sky are smooth in the ray, so each
    # precision's error is its stepping's, with no hashed star cell for a ray's last bits to land in or miss
    smooth_desc, cam = workload_scene(native)
    smooth_desc.stars = 0.0
    smooth = native.Scene(smooth_desc, cam)
    renderer.set_tracing("fp64", **EXTENDED_LAW)
    if not renderer.draw(smooth):
        raise RuntimeError("draw failed")
    ref64 = renderer.read_hdr().astype(np.float64)
    out["without_stars"] = {}
    for p in ("fp32", "mixed"):
        renderer.set_tracing(p, **EXTENDED_LAW)
        if not renderer.draw(smooth):
            raise RuntimeError("draw failed")
        out["without_stars"][p] = rel_error(renderer.read_hdr(), ref64)
    smooth.close()
    renderer.set_tracing("fp32")
    if own:
        renderer.close()
        scene.close()
    return out


def huge_frame(native, width=3840, height=2160, side=8, launch_tiles=1 << 20, draws=3):
    """The huge frame: 3840x2160 at 8x8 rays a pixel (530,841,600 rays) as a batch of one view in slices
    of 2^20 tiles (67 M rays, ~6 ms a launch: watchdog-sized; each pixel's 64 rays summed in one wave,
    under the default schedule), beside the renderer's banded launches of the tile kernel."""
    desc, cam = workload_scene(native, samples=side)
    out = {"width": width, "height": height, "rays_per_pixel": side * side, "rays": width * height * side * side,
           "launch_tiles": launch_tiles}
    # steps not counted, as the timed frames (the batch's steps equal the renderer's, tested)
    b = native.HoleBatch([(desc, cam, width, height)], native.tracing(precision="fp32", disc="painted", count_steps=False))
    b.set_schedule(0, 2, 64, 8, launch_tiles)
    ms = []
    for _ in range(draws + 1):
        b.draw()
        ms.append(b.stats().batchMs)
    st = b.stats()
    out["batch"] = {"ms": float(np.median(ms[1:])), "rays_per_s": st.rays / (np.median(ms[1:]) * 1e-3),
                    "launches": -(-(width * height) // launch_tiles)}  # a tile: one pixel's 64 rays
    b.close()
    r = native.Renderer(width, height)
    s = native.Scene(desc, cam)
    r.set_tracing("fp32", count_steps=False)
    tr = []
    for _ in range(draws + 1):
        if not r.draw(s):
            raise RuntimeError("draw failed")
        tr.append(r.hole_stats().traceMs)
    out["renderer_trace"] = {"ms": float(np.median(tr[1:])), "rays_per_s": out["rays"] / (np.median(tr[1:]) * 1e-3)}
    r.set_tracing("fp32")  # one counted frame for the steps
    if not r.draw(s):
        raise RuntimeError("draw failed")
    st = r.hole_stats()
    out["steps_per_ray"] = st.steps / st.rays
    r.close()
    s.close()
    log(f"huge frame: {out}")
    return out


def precision_study(native, width=3840, height=2160):
    """The precision study: 3840x2160, one ray a pixel, spins {0, .5, .9, .998} x inclinations {17, 45,
    75, 89}, eye at r 50, painted disc ISCO..20, Doppler 1: FP32, mixed and FP64 traces, each image's
    relative HDR error against its FP64 trace, and rays/s; then the same views made smooth in the ray
    (the analytic disc, no stars), each precision's error against FP64 again."""
    from artful_amd.workloads import view_of
    r = native.Renderer(width, height)
    scene = native.Scene()
    rows = []
    for spin in (0.0, .5, .9, .998):
        for incl in (17, 45, 75, 89):
            desc, cam, _, _ = view_of(spin, incl, 50.0, 20.0, width, height, fov_y=.35)
            scene.set_hole(desc)
            scene.set_camera(cam)
            row = {"spin": spin, "inclination": incl}
            hdr = {}
            for p in ("fp64", "fp32", "mixed"):
                r.set_tracing(p)
                ms = []
                for _ in range(3):
                    if not r.draw(scene):
                        raise RuntimeError("draw failed")
                    ms.append(r.hole_stats().traceMs)
                hdr[p] = r.read_hdr()
                row[f"{p}_rays_per_s"] = width * height / (min(ms) * 1e-3)
            ref = hdr["fp64"].astype(np.float64)
            for p in ("fp32", "mixed"):
                e = rel_error(hdr[p], ref)
                rel = np.abs(hdr[p].astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-3)
                e["p99.9"] = float(np.quantile(rel, .999))
                row[f"{p}_error"] = e
            row["mixed_over_fp64"] = row["mixed_rays_per_s"] / row["fp64_rays_per_s"]
            rows.append(row)
            log(f"precision study: {row}")
    # the same 16 views made smooth in the ray (the analytic r^-3 disc, the stars off): each precision's
    # error is then its stepping's alone, with no texel row or hashed star cell for a ray's last bits to pick
    smooth = []
    for spin in (0.0, .5, .9, .998):
        for incl in (17, 45, 75, 89):
            desc, cam, _, _ = view_of(spin, incl, 50.0, 20.0, width, height, fov_y=.35)
            desc.stars = 0.0
            scene.set_hole(desc)
            scene.set_camera(cam)
            hdr = {}
            for p in ("fp64", "fp32", "mixed"):
                r.set_tracing(p, disc="analytic", emissivity=3.0)
                if not r.draw(scene):
                    raise RuntimeError("draw failed")
                hdr[p] = r.read_hdr()
            ref = hdr["fp64"].astype(np.float64)
            smooth.append({"spin": spin, "inclination": incl, **{f"{p}_error": rel_error(hdr[p], ref) for p in ("fp32", "mixed")}})
    r.close()
    scene.close()
    summary = {p: {"max_of_max": max(x[f"{p}_error"]["max"] for x in rows), "max_of_mean": max(x[f"{p}_error"]["mean"] for x in rows),
                   "max_share_over_1e-3": max(x[f"{p}_error"]["share_over_1e-3"] for x in rows),
                   "smooth_max_of_mean": max(x[f"{p}_error"]["mean"] for x in smooth),
                   "smooth_max_share_over_1e-3": max(x[f"{p}_error"]["share_over_1e-3"] for x in smooth)} for p in ("fp32", "mixed")}
    return {"width": width, "height": height, "images": rows, "smooth_images": smooth, "summary": summary}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50, help="untimed frames first (50: ~60 ms, the GPU's clock up to speed)")
    ap.add_argument("--settle", type=float, default=0.3,
                    help="seconds of untimed frames before the warm-up frames (the frame path's first frames after idle run slow)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="only the FP32 frames and the FP32 kernel timing (for rocprofv3 runs)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-views leg")
    ap.add_argument("--no-study", action="store_true", help="skip the 4K precision study")
    ap.add_argument("--only", choices=["batch", "study", "adjoint", "animation", "huge", "extended", "render"], help="run one extra leg alone and print its JSON")
    ap.add_argument("--batch-schedule", help="workgroups,chunk,refill,stride for the batch leg instead of tuning (0: default; workgroups -1: a workgroup a tile)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead: a plain device-to-device copy of this many bytes, 5 times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the tracer runs on one device (sharding a frame by row bands or a batch by "
              "views across GPUs is not built). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from artful_amd import native
    if native.device_count() < 1:
        raise SystemExit("no HIP device: bench.py needs an MI355X")
    if args.copy_bytes:
        L = native.lib()
        a, b = L.artfHipMalloc(args.copy_bytes), L.artfHipMalloc(args.copy_bytes)
        for _ in range(5):
            if L.artfHipCopy(b, a, args.copy_bytes, None) or L.artfHipStreamSync(None):
                raise RuntimeError("copy failed")
        L.artfHipFree(a), L.artfHipFree(b)
        print(json.dumps({"copy_bytes": args.copy_bytes, "launches": 5}))
        return
    if args.only:
        sched = tuple(int(v) for v in args.batch_schedule.split(",")) if args.batch_schedule else None
        legs = {"batch": lambda n: batched_views(n, schedule=sched), "study": precision_study, "adjoint": adjoint_fit,
                "animation": animation, "huge": huge_frame, "extended": extended_form, "render": render_scene}
        leg = legs[args.only](native)
        print(json.dumps({args.only: leg}), flush=True)
        return
    desc, cam = workload_scene(native)
    scene = native.Scene(desc, cam)
    renderer = native.Renderer(WIDTH, HEIGHT)
    log(renderer.description())

    # the dominant kernel alone, timed with HIP events on its stream: the roofline's and the tracers' rates
    # (first, so the frames below start with the GPU's clock up from the idle of setting up)
    kernel = None if args.profile else {p: kernel_ms(native, scene, p) for p in ("fp32", "fp32_fast", "mixed", "fp64")}
    log(f"kernel ms: {kernel}")
    fp32 = time_frames(native, renderer, scene, "fp32", args.steps, args.warmup, args.settle)
    rays = fp32["rays"]
    value = rays * args.steps / fp32["wall_s"]
    log(f"fp32: {fp32}")
    if args.profile:
        ms = kernel_ms(native, scene, "fp32")
        print(json.dumps({"rays_per_s": value, "ms_per_step": fp32["wall_s"] / args.steps * 1e3, "trace_kernel_ms": ms,
                          "gpu_span_ms_per_step": fp32["gpu_span_ms"] / args.steps, "frame_trace_ms": fp32["trace_ms"], "post_ms": fp32["post_ms"],
                          "achieved_tflops": fp32["steps_per_frame"] * FLOPS_PER_STEP[SPIN != 0] / (ms * 1e-3) / 1e12}))
        return
    hdr32 = renderer.read_hdr()

    others = {}
    hdrs = {}
    for p in ("fp32_fast", "mixed", "fp64"):
        others[p] = time_frames(native, renderer, scene, p, max(3, args.steps // 4), 1, args.settle)
        hdrs[p] = renderer.read_hdr()
        log(f"{p}: {others[p]}")
    ref64 = hdrs["fp64"].astype(np.float64)
    errors = {p: rel_error(h, ref64) for p, h in (("fp32", hdr32), ("fp32_fast", hdrs["fp32_fast"]), ("mixed", hdrs["mixed"]))}

    extended = extended_form(native, renderer, scene)
    log(f"extended: {extended}")

    flops = fp32["steps_per_frame"] * FLOPS_PER_STEP[SPIN != 0]
    achieved = flops / (kernel["fp32"] * 1e-3) / 1e12
    traffic = traffic_from_profiles()
    cpu = None if args.no_cpu else cpu_baseline(native)
    if cpu is not None and cpu["kind"] == "reference":
        # the CPU frame traces the GPU frame's rays (its HDR is the GPU's bit for bit), so its steps are the
        # frame's, which equal the oracle's count of the same rays (tests/golden/steps.json)
        cpu["steps_per_ray"] = fp32["steps_per_frame"] / rays
    adjoint = adjoint_fit(native)
    log(f"adjoint: {adjoint}")
    renderer.close()
    scene.close()
    anim = animation(native)
    published = render_scene(native)
    huge = huge_frame(native)
    batch = None if args.no_batch else batched_views(native)
    study = None if args.no_study else precision_study(native)

    line = {
        "metric": "Kerr rays/s at 1080p (FP32 vs FP64 plain, <=1e-3 rel err); adjoint/forward time",
        "value": value, "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": fp32["wall_s"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "gpu_span_ms_per_step": fp32["gpu_span_ms"] / args.steps, "settle": fp32["settle"],
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "kerr-1080p-frame", "width": WIDTH, "height": HEIGHT, "rays_per_pixel": SAMPLES * SAMPLES,
                   "spin": SPIN, "observer_r": RADIUS, "inclination_deg": INCLINATION, "disc": "painted, ISCO..30M",
                   "doppler": 1.0, "bloom_levels": fp32["bloom_levels"], "parallelism": "single-gpu"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic},
        "cpu_baseline": cpu,
        "trace_rays_per_s": {p: rays / (ms * 1e-3) for p, ms in kernel.items()},
        "fp32_over_fp64": kernel["fp64"] / kernel["fp32"],
        "trace_kernel_ms": kernel,
        "frame_stages_ms": {"trace": fp32["trace_ms"], "post": fp32["post_ms"]},
        "steps_per_ray": fp32["steps_per_frame"] / rays, "active_lane_fraction": fp32["active_lanes"],
        "rel_error_vs_fp64": errors,
        "extended_form": extended,
        "adjoint_over_forward": adjoint["adjoint_over_forward"],
        "adjoint": adjoint,
        "animation": anim,
        "render_scene": published,
        "huge_frame": huge,
        "batched_views": batch,
        "precision_study": study,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
