"""View sets for the batched tracer (include/artful/batch.h).

uneven_views: views of uneven size and scene, each drawn from its own seeded stream: width and
height uniform in 128..1024, spin U[0, .998], inclination U[5, 85] degrees from the disc's axis, eye
at r U[20, 100] looking at the hole, disc from the last stable orbit out to U[10, 40]; Doppler 1, one
ray a pixel. The eye sits at r (0, cos i, sin i) about the +y axis, up +y (or -z once its elevation
passes 1 rad), as the reference's test scenes place it (tests/tests.c:674-687); the field of view
(fov_y radians) is the same for all."""
from __future__ import annotations

import math

import numpy as np

from . import native


def view_of(spin, inclination_deg, radius, outer, width, height, fov_y=.6, samples=1):
    desc = native.hole_desc(spin=float(spin), axis=(0, 1, 0), discInner=0.0, discOuter=float(outer), doppler=1.0, samples=samples)
    i = math.radians(float(inclination_deg))
    up = (0, 1, 0) if math.pi / 2 - i <= 1 else (0, 0, -1)
    cam = native.camera((0.0, radius * math.cos(i), radius * math.sin(i)), (0, 0, 0), up, fov_y, .1)
    return desc, cam, int(width), int(height)


def uneven_views(count=64, seed=4, sizes=(128, 1024), fov_y=.6, samples=1):
    """[(desc, camera, width, height)] and the parameters drawn, one dict a view."""
    rng = np.random.default_rng(seed)
    views, params = [], []
    for _ in range(count):
        w, h = (int(x) for x in rng.integers(sizes[0], sizes[1] + 1, 2))
        p = dict(width=w, height=h, spin=float(rng.uniform(0, .998)), inclination=float(rng.uniform(5, 85)),
                 radius=float(rng.uniform(20, 100)), outer=float(rng.uniform(10, 40)))
        views.append(view_of(p["spin"], p["inclination"], p["radius"], p["outer"], w, h, fov_y, samples))
        params.append(p)
    return views, params
