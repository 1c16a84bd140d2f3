import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "friendly-journey_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


# torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the system
# one libvxl links: whichever loads first serves both. Load torch's first, as bench.py
# does, so torch's device calls work in tests that hand torch buffers to libvxl.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def ctx():
    import vxl
    c = vxl.Context(0)
    yield c
    c.close()
