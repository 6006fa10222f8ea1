import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_NAMES = ["sparse_directed", "dense_directed", "undirected", "components"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden():
    return {name: load_golden(name) for name in GOLDEN_NAMES}


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; fails (does not skip) when absent, so a GPU run never
    passes on a silent fallback."""
    import torch

    import meerkat_amd

    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    meerkat_amd.lib()  # raises if libmeerkat_hip.so is missing
    return torch.device("cuda:0")
