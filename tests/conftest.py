"""Shared test setup: import paths, the gpu marker, and small scenes."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

for path in (ROOT / "monki-gpu_amd", ROOT, ROOT / "tests"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libmonki_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU reference runs")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle

    return Oracle()
