import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    here = os.path.join(ROOT, "tests", "golden")

    def load(name):
        return np.load(os.path.join(here, name + ".npz"))
    return load


def rel_l2(a, b):
    import torch
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-300))


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU in this environment")
    import inc_amd._lib as L
    L.load()  # raises if the HIP library is missing: GPU tests never run without it
    return torch.device("cuda")
