import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ngp-encode-server_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def nes():
    import nes_amd as N
    N.lib()
    return N


@pytest.fixture(scope="session")
def scene1(oracle):
    """The seeded scene of the render configs (seed 1)."""
    return oracle.scene(1)


@pytest.fixture(scope="session")
def gpu_ctx(nes, scene1):
    t, w, o = scene1
    ctx = nes.Context(t, w, o)
    yield ctx
    ctx.close()
