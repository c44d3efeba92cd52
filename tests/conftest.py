import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


class Fixture:
    """The reference's shipped output file, extracted to .npz (tests/golden)."""

    def __init__(self, name):
        self._z = np.load(os.path.join(GOLDEN, name))

    def __getitem__(self, path):
        return self._z[path.strip("/").replace("/", "|")]

    def __contains__(self, path):
        return path.strip("/").replace("/", "|") in self._z


@pytest.fixture(scope="session")
def fixture_small():
    return Fixture("tng300_mock_small.npz")


@pytest.fixture(scope="session")
def fixture_large():
    return Fixture("tng300_mock_large.npz")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
