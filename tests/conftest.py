"""pytest configuration: the `gpu` marker and shared helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libsimd2.so")


def load_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_srgemm():
    return load_npz("host_srgemm_cases.npz")


@pytest.fixture(scope="session")
def golden_converge():
    return load_npz("converge_cases.npz")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """torch on the GPU plus the loaded HIP library; fails loudly without them."""
    import torch
    import simd2_amd
    assert torch.cuda.is_available(), "GPU test needs a visible GPU"
    simd2_amd.lib()
    return torch
