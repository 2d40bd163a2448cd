import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "neon-physics-fluxion_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfluxion_hip.so)")


@pytest.fixture(scope="session")
def ctx():
    import fluxion as fxn
    c = fxn.GpuContext.new(0)
    print("device:", c.adapter_name, flush=True)
    yield c
    c.close()
