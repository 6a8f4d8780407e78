import os
import sys

import pytest

# torch ships its own HIP runtime under the same soname as the system one
# that libscl.so links.  Importing torch before libscl.so is loaded makes the
# process use a single runtime (torch's), which libscl.so then binds to;
# loading libscl.so first would leave torch with a second runtime that
# cannot see the GPU.
try:
    import torch  # noqa: F401
except Exception:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import spectral3d

    return spectral3d
