import os
import sys

import pytest

# Torch ships its own HIP runtime and links it by the unversioned name libamdhip64.so.  Loaded
# after libpartia_mi355x.so (which links /opt/rocm's libamdhip64.so.7), it would be a second
# HIP runtime in the process and find no device; loaded first, its runtime also serves the
# library (same SONAME).  So the tests import torch, when present, before the library.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpartia_mi355x.so")


@pytest.fixture(scope="session")
def ctx():
    from distorteddustydiscanalysistoolkits_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()
