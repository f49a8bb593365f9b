import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def handle():
    import torch
    import psd_projection_amd as P
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = P.Handle(0, seed=1234)
    yield h
    h.close()
