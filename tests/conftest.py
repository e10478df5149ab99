import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

# One HIP runtime per process: torch's bundled libamdhip64.so.7 must be loaded
# before the product library (same soname), or torch finds no device.
try:
    import torch  # noqa: F401,E402
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X; runs the HIP path through the C ABI")
