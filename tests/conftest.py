import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7) as
# the system one librf_hip.so links.  Loading torch first makes librf_hip.so
# bind to that already-loaded runtime, so one process never holds two HIP
# runtimes (the GPU tests use torch tensors as device buffers).
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import randomfield_amd as R
    c = R.Context(0)
    yield c
    c.close()
