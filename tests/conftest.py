import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "halo2-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def co():
    from oracle import coracle

    coracle.build()
    return coracle


@pytest.fixture(scope="session")
def po():
    from oracle import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def lib():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    from halo2_amd import _lib

    return _lib.load()


def rand_fr(co, n: int, seed: int) -> np.ndarray:
    """Seeded uniform Fr (Montgomery limbs) from the oracle's ChaCha20 stream."""
    return co.chacha_fr(seed, 0, n)
