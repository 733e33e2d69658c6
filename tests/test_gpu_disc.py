"""The disc map painted on the GPU (include/artful/device_disc.h) against the reference's map: the
gas, the disc's dust, the sky and, with the host's colour and ring tables, every table, bit for bit
(the reference's digests of each golden scene's map)."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from artful_amd import abi
from tests.golden import cases

pytestmark = pytest.mark.gpu


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


@pytest.mark.parametrize("name", ["config1", "defaults", "inner_given", "shadow_r14", "grid_a0.998_i89_d1"])
def test_device_map_equals_reference(name, gpu, golden):
    g = golden["scenes"][name]
    c = g["clamped"]
    m = gpu.device_map(c["spin"], c["discInner"], c["discOuter"])
    assert digest(m[:abi.ROWS_AT]) == g["map_gas"]
    assert digest(m[abi.ROWS_AT:]) == g["map_tables"]


def test_scene_with_its_disc_on_the_device(gpu):
    """The scene keeps no host map or fluid; its version counters still move as the host scene's."""
    desc, cam = cases.make(cases.CASES["config1"])
    s = gpu.Scene(cam=cam)
    s.set_disc_on_device(0)
    s.set_hole(desc)
    assert s.map() is None and s.fluid()[0] is None
    v0 = s.map_version()
    desc.spin = .5
    s.set_hole(desc)
    assert s.map_version() == v0 + 1
    s.close()
