"""tests/golden/steps.json against the oracle and itself, on the CPU: the configurations the oracle traces
in a moment (config 1, 320x180; config 5, the adjoint's 512x512 view) counted again must give the same
steps, lane-slots and histogram; every configuration's histogram must add up to its rays and steps; and
the lane-slot rule (each tile tracer wave's longest ray times its rays) must agree with a plain loop over
the waves."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from tests.golden import make_steps as M

STEPS = json.loads((Path(__file__).parent / "golden" / "steps.json").read_text())


def entries():
    for name, v in STEPS.items():
        if "histogram" in v:
            yield name, v
        else:
            for view, w in v.items():
                yield f"{name} {view}", w


@pytest.mark.parametrize("name", ["config1", "config5"])
def test_counted_again(name):
    assert M.MAKERS[name]() == STEPS[name]


def test_histograms_add_up():
    for name, v in entries():
        h = np.array(v["histogram"], np.int64)
        assert h[:, 1].sum() == v["rays"] and (h[:, 0] * h[:, 1]).sum() == v["steps"], name
        assert v["steps"] <= v.get("lane_slots", v["steps"]) <= h[:, 0].max() * v["rays"], name
    assert sum(STEPS["config4"]["views"]) == STEPS["config4"]["steps"]


@pytest.mark.parametrize("side", [1, 2, 3, 4, 5, 6, 8])
def test_lane_slot_rule(side):
    steps = np.random.default_rng(side).integers(1, 300, (side * side, 37, 53)).astype(np.uint32)
    assert M.lane_slots_fast(steps) == M.lane_slots(steps)
