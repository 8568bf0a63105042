"""Shared test configuration: the `gpu` marker and import paths."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    return torch


@pytest.fixture(autouse=True)
def _release_workspace_pins():
    """Graphs a test captures die with the test; the wavefront workspace pins they took are
    released after it, so a later test may grow the workspace."""
    yield
    from lsystemeditor_amd import abi
    if abi._lib is not None:
        n = abi._lib.pathTracerWavefrontPins()
        if n:
            abi._lib.pathTracerWavefrontUnpin(n)
