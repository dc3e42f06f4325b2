import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libxrp on the device)")


@pytest.fixture(scope="session")
def ctx():
    import xray_projection_render_amd as X
    c = X.Context(0)
    yield c
    c.close()
