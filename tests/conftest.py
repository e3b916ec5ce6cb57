import os
import sys

import pytest

# torch (device buffers in a few GPU tests) loads its own HIP runtime; importing it
# before the library is first loaded keeps one runtime in the process, as bench.py does
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "axia-fem_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import axia_gpu
    c = axia_gpu.Context(0)
    yield c
    c.close()
