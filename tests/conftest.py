"""pytest configuration: markers and shared fixtures.

`gpu` tests need a visible GPU and the built libmf_amd.so; everything else
runs on the CPU (oracle against golden vectors and closed forms, the host
build of the fit core against the oracle, the C ABI's symbol table).
"""
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))

from oracle_binding import load_pkg  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test selected but no GPU is visible")
    return torch.device("cuda", 0)
