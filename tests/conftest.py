import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkq on the device)")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


def _emit(msg: str) -> None:
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any device work, so a failing run still says which case it was
    if request.node.get_closest_marker("gpu") is not None:
        _emit(f"[gpu-case] {request.node.nodeid}")
    yield
