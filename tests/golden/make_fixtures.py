"""Writes tests/golden/fixtures.json (and small .npz arrays) from the reference's own library,
oracle/_ref/libartful_ref.so, which oracle/build.py compiles from the reference's sources.

Every value here is an output of the reference's code on stated inputs: the disc map and fluid
(src/scene/hole.c, fluid.c), the frame uniforms (hole.c:273-371), the tracer's HDR through the
AVX2 and portable kernels (lens.inc via kernels*.c), whole software frames (BGRA8), the last stable
orbit and orbit speeds. Large outputs are stored as SHA-256 digests of their bytes; a few small HDR
frames are stored whole.

    python tests/golden/make_fixtures.py              (everything)
    python tests/golden/make_fixtures.py --layouts    (only the struct layouts and enum values)
"""
from __future__ import annotations

import hashlib
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from artful_amd import abi  # noqa: E402  (layouts only)
from tests.golden import cases  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(a: np.ndarray) -> str:
    """SHA-256 of an array's bytes, first 16 hex digits."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def layouts(L) -> dict:
    """sizeof and offsetof of the public structs, and the last value of each public enum, as the
    reference's headers give them (oracle/ref/ref_driver.c's table)."""
    out: dict = {}
    for i in range(L.refLayoutCount()):
        out.setdefault(L.refLayoutType(i).decode(), {})[L.refLayoutField(i).decode() or "sizeof"] = int(L.refLayoutAt(i))
    out["enums"] = {"ARTF_BACKEND_METAL": L.refEnumValue(0), "ARTF_SURFACE_METAL": L.refEnumValue(1), "ARTF_SIMD_256": L.refEnumValue(2)}
    return out


def main() -> None:
    if "--layouts" in sys.argv:  # only the layouts, into the fixtures already written
        fx = json.loads((HERE / "fixtures.json").read_text())
        fx["layouts"] = layouts(O.reference())
        (HERE / "fixtures.json").write_text(json.dumps(fx, indent=1, sort_keys=True))
        print("wrote the layouts into", HERE / "fixtures.json")
        return
    if not O.reference_available():
        raise SystemExit("oracle/_ref/libartful_ref.so is not built (python oracle/build.py, with the reference tree present)")
    L = O.reference()
    assert L.refMapFloats() == abi.MAP_FLOATS and L.refHoleFrameBytes() == 368
    out: dict = {"isco": {}, "orbit_speed": {}, "scenes": {}, "fluid_alone": {}}
    for a in cases.SPINS_ISCO:
        out["isco"][repr(a)] = float(np.float32(L.refIsco(a)))
    for a, r in cases.ORBITS:
        out["orbit_speed"][f"{a!r},{r!r}"] = float(np.float32(L.refOrbitSpeed(a, r)))
    d = abi.BlackHoleDesc()
    L.refDefaults(d)
    out["defaults"] = {k: (getattr(d, k).tolist() if hasattr(getattr(d, k), "tolist") else getattr(d, k)) for k, _ in abi.BlackHoleDesc._fields_}
    arrays = {}
    for name, case in cases.CASES.items():
        desc, cam = cases.make(case)
        s = O.RefScene(desc, cam)
        m, fl = s.map(), s.fluid()
        entry = {"map_gas": digest(m[:abi.ROWS_AT]), "map_tables": digest(m[abi.ROWS_AT:]), "fluid": digest(fl),
                 "fluid_time": L.refFluidTime(s.ptr), "frames": {}, "lens": {}, "bgra": {}}
        clamped = abi.BlackHoleDesc()
        L.refSceneDesc(s.ptr, clamped)
        entry["clamped"] = {"spin": clamped.spin, "discInner": clamped.discInner, "discOuter": clamped.discOuter,
                            "discHeight": clamped.discHeight, "samples": clamped.samples}
        for (w, h) in case["sizes"]:
            f, n = s.frame(w, h)
            entry["frames"][f"{w}x{h}"] = {"bytes": bytes(f).hex(), "bloom_pixels": n}
            hdr = s.lens(w, h, 2)
            hdr_portable = s.lens(w, h, 0)
            assert np.array_equal(hdr.view(np.uint32), hdr_portable.view(np.uint32)), "AVX2 and portable kernels differ"
            entry["lens"][f"{w}x{h}"] = digest(hdr)
            entry["bgra"][f"{w}x{h}"] = digest(s.draw(w, h, 0, 1))
            if case.get("store") == (w, h):
                arrays[f"{name}_{w}x{h}_hdr"] = hdr
        if case.get("advance"):
            desc2, _ = cases.make(case)
            desc2.time = case["advance"]
            s.set(desc2)
            entry["advanced"] = {"time": case["advance"], "fluid": digest(s.fluid()), "fluid_time": L.refFluidTime(s.ptr),
                                 "map_tables": digest(s.map()[abi.ROWS_AT:])}
            w, h = case["sizes"][0]
            entry["advanced"]["lens"] = digest(s.lens(w, h, 2))
        s.close()
        out["scenes"][name] = entry
        print(name, "done", flush=True)
    for inner, outer, spin, times in cases.FLUIDS:
        f = L.refFluidNew(inner, outer, spin, 0.0)
        hs = []  # in the order the times were reached: a list, since the order matters
        for t in times:
            L.refFluidAdvance(f, t)
            hs.append([t, digest(np.ctypeslib.as_array(L.refFluidField(f), shape=(abi.FLUID_WIDTH * abi.FLUID_HEIGHT,)).copy())])
        L.refFluidFree(f)
        out["fluid_alone"][f"{inner!r},{outer!r},{spin!r}"] = hs
    out["layouts"] = layouts(L)
    (HERE / "fixtures.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    np.savez_compressed(HERE / "hdr_small.npz", **arrays)
    print("wrote", HERE / "fixtures.json", HERE / "hdr_small.npz")


if __name__ == "__main__":
    main()
