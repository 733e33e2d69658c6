"""The scenes golden vectors are taken on: shared by make_fixtures.py (which runs the reference's
library) and the tests (which run this project's code and the oracle on the same inputs).

Scenes follow the reference's test helper holeScene (tests/tests.c:674-687): axis +y, the eye at
distance r and elevation `tilt` above the disc's plane, looking at the hole, up +y (or -z when the
elevation passes 1 rad)."""
from __future__ import annotations

import math

from artful_amd import abi

SPINS_ISCO = [0.0, .3, .5, .6, .7, .9, .95, .998, -.5, -.998]
ORBITS = [(0.0, 6.0), (.6, 10.0), (.9, 3.0), (.998, 2.0), (.5, 40.0)]
# (inner, outer, spin, times advanced to, in order)
FLUIDS = [(3.83, 34.0, .6, [0.0, 1.0, 2.0, 2.05, 12.0])]


def holeish(distance, tilt, fov=.9, **kw):
    return dict(distance=distance, tilt=tilt, fov=fov, desc=kw)


def _grid():
    out = {}
    for spin in (0.0, .6, .9, .998):
        for incl in (17, 60, 89):
            for dop in (0.0, 1.0):
                name = f"grid_a{spin}_i{incl}_d{int(dop)}"
                out[name] = dict(distance=50.0, tilt=math.radians(90 - incl), fov=.35,
                                 desc=dict(spin=spin, discOuter=20.0, doppler=dop), sizes=[(48, 27)])
    return out


CASES = {
    # the CPU-sized view: 320x180, spin .9, r 50, inclination 80, fovY .35, disc ISCO..20, Doppler 1
    "config1": dict(distance=50.0, tilt=math.radians(10), fov=.35, desc=dict(spin=.9, discOuter=20.0, doppler=1.0),
                    sizes=[(320, 180), (64, 36)], store=(64, 36)),
    # the 1080p workload's view at a tenth of the size: spin .998, r 60, inclination 60, disc ISCO..30, 2x2 rays
    "config2_small": dict(distance=60.0, tilt=math.radians(30), fov=.35, desc=dict(spin=.998, discOuter=30.0, doppler=1.0, samples=2),
                          sizes=[(192, 108)]),
    # tests.c:4317-4340: spin .9, bloom 1, 2x2 rays, shutter 3, a moving eye
    "moving_eye": dict(distance=26.0, tilt=.06, fov=.9, desc=dict(spin=.9, doppler=.5, stars=1.0, bloom=1.0, samples=2, shutter=3.0,
                                                                velocity=(.3, 0, 0)),
                       sizes=[(320, 200), (101, 67), (3, 2), (1, 1)], store=(101, 67)),
    # tests.c:4402-4420: the scene the reference holds its GPU against
    "gpu_vs_cpu": dict(distance=26.0, tilt=.06, fov=.9, desc=dict(spin=.8, doppler=.6, bloom=1.0, samples=2, shutter=2.0,
                                                                velocity=(.2, 0, -.1), time=40.0),
                       sizes=[(320, 200)]),
    # tests.c:4249-4275: Schwarzschild shadow at r = 14, the disc pushed out, seen along the axis
    "shadow_r14": dict(distance=14.0, tilt=1.57079633, fov=2.6 * math.asin(math.sqrt(27) * math.sqrt(1 - 2 / 14) / 14),
                       desc=dict(spin=0.0, discInner=10.0, discOuter=60.0, discGain=200.0, samples=4, stars=0.0, bloom=0.0),
                       sizes=[(160, 100)]),
    # tests.c:4277-4315: Kerr shadow edges lit by stars alone
    "kerr_edges_a95": dict(distance=60.0, tilt=0.0, fov=.36, desc=dict(spin=.95, discGain=0.0, discThickness=0.0, stars=1e4, bloom=0.0),
                           sizes=[(320, 200)]),
    # the reference's defaults (axis (0,-1,0), spin .6, disc ISCO..34), a given inner edge, and 8x8 rays
    "defaults": dict(distance=30.0, tilt=.2, fov=.9, desc=dict(), axis_default=True, sizes=[(80, 50)]),
    "inner_given": dict(distance=40.0, tilt=.3, fov=.6, desc=dict(spin=-.5, discInner=5.0, discOuter=12.0, samples=9, doppler=1.0),
                        sizes=[(16, 9)]),
    # time moves the gas: the fluid advanced to t = 40
    "advance": dict(distance=26.0, tilt=.06, fov=.9, desc=dict(spin=.6, doppler=0.0, stars=0.0, bloom=0.0), sizes=[(64, 40)], advance=40.0),
    **_grid(),
}


def make(case: dict):
    """(ArtfBlackHoleDesc, ArtfCamera) for a case, built from the reference's defaults
    (artfBlackHoleDefaults, src/scene/hole.c:225-240) restated here so both libraries get the
    same bytes."""
    d = abi.BlackHoleDesc()
    d.axis = abi.Vec3(0, -1, 0)
    d.mass, d.spin, d.discOuter, d.discHeight, d.discThickness, d.discGain = 1.0, .6, 34.0, .03, 1.0, 2.1
    d.stars, d.bloom, d.exposure, d.samples = .4, 4.5, 1.0, 1
    if not case.get("axis_default"):
        d.axis = abi.Vec3(0, 1, 0)
        d.spin, d.bloom = 0.0, 0.0  # holeScene's own changes (tests.c:676-682)
    for k, v in case["desc"].items():
        setattr(d, k, abi.Vec3(*v) if k in ("velocity", "position", "axis") else v)
    r, tilt = case["distance"], case["tilt"]
    pos = abi.Vec3(0.0, math.sin(tilt) * r, math.cos(tilt) * r)
    up = abi.Vec3(0, 1, 0) if abs(tilt) < 1.0 else abi.Vec3(0, 0, -1)
    cam = abi.Camera(pos, abi.Vec3(0, 0, 0), up, case["fov"], .1)
    return d, cam
