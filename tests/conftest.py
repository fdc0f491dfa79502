import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libdenrim_hip.so")


@pytest.fixture(scope="session")
def renderer():
    # No skip: a GPU test that cannot reach the device or the library fails.
    from denrimrendererkit_amd import DenrimRenderer
    return DenrimRenderer(0)
