import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the reference tree to build oracle/_ref")


@pytest.fixture(scope="session")
def gpu_lib():
    """The gfx950 library with the golden CIA table registered."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from radtrancode_amd import cirsradg
    from tests.golden.golden_io import load_cia

    cirsradg.lib()
    cirsradg.register_cia(load_cia())
    return cirsradg
