import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("lsmpqs-1.0_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

# Some GPU tests hand torch-allocated device buffers to the library.  torch
# ships its own HIP runtime; importing it before the library is first loaded
# makes the library bind to that same runtime (one runtime per process).  The
# other order leaves two runtimes, and torch's then finds no device.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False
