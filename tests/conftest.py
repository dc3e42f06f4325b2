import os
import sys

import pytest
# torch ships HIP and HSA runtimes with the same sonames as the system ROCm
# that libxrp.so links (libamdhip64.so.7, libhsa-runtime64.so.1); the first
# loaded serves the whole process, and torch initialises only on its own.
# Loading torch before libxrp gives every test the runtime bench.py uses.
import torch  # noqa: F401,E402

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
