"""Test setup: the `gpu` marker (tests that need an MI355X), and shared fixtures.

CPU tests check the oracle against the reference's golden vectors and the mathematics, the host C
library (scene, fluid, frame) against the same vectors, and that the C ABI library loads and exports
what include/*.h declares. GPU tests compare the HIP path with the oracle through the C ABI."""
from __future__ import annotations

import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def golden():
    return json.loads((ROOT / "tests" / "golden" / "fixtures.json").read_text())


@pytest.fixture(scope="session")
def native():
    from artful_amd import native as n
    n.lib()
    return n


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.oracle()
    return o


_scenes: dict = {}


@pytest.fixture(scope="session")
def scene_of(native):
    """A product scene per golden case, built once (the disc map takes a while)."""
    from tests.golden import cases

    def get(name):
        if name not in _scenes:
            desc, cam = cases.make(cases.CASES[name])
            _scenes[name] = native.Scene(desc, cam)
        return _scenes[name]
    return get


@pytest.fixture(scope="session")
def gpu(native):
    if native.device_count() < 1:
        pytest.fail("no HIP device: the gpu tests need an MI355X")
    return native
